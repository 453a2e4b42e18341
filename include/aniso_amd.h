/* C ABI of the MI355X Gaussian-splat rasterizer (libaniso_amd.so, HIP kernels for gfx950).
 *
 * This is the drop-in boundary for the reference's GPU renderer. Every entry point takes plain
 * pointers and sizes; none takes a C++ or torch type. Status: 0 = ok, nonzero = error, with the
 * message in ans_last_error() (thread-local). The C++ shim in aniso_amd/gpu_render.hpp rethrows
 * errors as std::runtime_error, the reference's convention (gpu_render.cu:18-21).
 *
 * Threading: one HIP stream per renderer handle; calls on one handle are not thread-safe;
 * distinct handles may be used concurrently.
 */
#ifndef ANISO_AMD_H
#define ANISO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANS_OK 0
#define ANS_ERR_ARGUMENT 1
#define ANS_ERR_HIP 2
#define ANS_ERR_CAPACITY 3
#define ANS_ERR_NO_DEVICE 4

/* Pinhole camera, COLMAP convention: X_cam = R * X_world + t, R row-major, +x right, +y down,
 * +z forward. Doubles, like aniso::Camera (include/aniso/camera.h:16-26), so the host derives
 * the FP32 per-frame constants identically for the kernels and the CPU oracle. */
typedef struct ans_camera {
    double R[9];
    double t[3];
    double fx, fy, cx, cy;
    int32_t width, height;
} ans_camera;

/* Render options: aniso::RenderOptions (include/aniso/render.h:28-33) plus the tile shape and
 * the blend's LDS batch depth. Zero for tile_w / tile_h / batch_depth picks the defaults
 * (16 x 16, 256). Supported tiles: 8x8, 16x16, 32x16 (w x h). */
typedef struct ans_options {
    int32_t sh_degree;   /* clamped to the scene's degree */
    float splat_scale;   /* multiplies every standard deviation */
    int32_t tile_w, tile_h;
    int32_t batch_depth; /* splats staged in LDS per blend round: 64, 128 or 256 */
    int32_t flags;       /* ANS_FLAG_* */
} ans_options;

/* Blend every staged splat for every 8x8 pixel block, without the conservative per-block
 * culling. The output is identical; used to validate the culling. */
#define ANS_FLAG_NO_CULL 1
/* Batched views only: launch the batch's kernels directly instead of replaying its captured
 * HIP graph (same results; for comparison). */
#define ANS_FLAG_NO_GRAPH 2

/* aniso::RenderStats (include/aniso/render.h:9-13) plus the pair-buffer bookkeeping. Times are
 * GPU time per stage from the device wall clock, read by the frame's own kernels: project =
 * preprocess + depth sort, sort = pair expansion + pair sort + tile ranges, blend = blend +
 * output. */
typedef struct ans_stats {
    int64_t visible;
    int64_t tile_pairs;
    double project_ms, sort_ms, blend_ms;
    int64_t pair_capacity; /* pair slots allocated after this call */
    int32_t retries;       /* renders repeated because the pair count exceeded the capacity */
    int32_t reserved;
} ans_stats;

typedef struct ans_renderer ans_renderer;

/* 1 if a HIP device is present and usable (GpuRenderer::available, gpu_render.cu:262-265). */
int ans_available(void);

/* Message of the last error on this thread, "" if none. */
const char* ans_last_error(void);

/* Uploads an activated scene once (GpuRenderer::GpuRenderer(const Scene&), gpu_render.cu:267-292):
 * pos_xyz [n][3], scale_xyz [n][3] standard deviations, rot_wxyz [n][4] unit quaternions (w first),
 * opacity [n] in [0, 1], sh [n][(sh_degree+1)^2][3] laid out [coefficient][channel]
 * (include/aniso/gaussian.h:33-39). Host arrays are copied; the caller keeps ownership. */
int ans_renderer_create(int64_t n, int32_t sh_degree, const float* pos_xyz, const float* scale_xyz,
                        const float* rot_wxyz, const float* opacity, const float* sh, ans_renderer** out);

/* Same, from the raw values a 3DGS PLY stores: log-scales, opacity logits, unnormalized
 * quaternions. Activates on the host exactly as the reference loader does
 * (src/ply_loader.cpp:124-139: exp, 1/(1+exp(-x)), q/|q| in float). */
int ans_renderer_create_raw(int64_t n, int32_t sh_degree, const float* pos_xyz, const float* log_scale_xyz,
                            const float* quat_wxyz, const float* opacity_logit, const float* sh,
                            ans_renderer** out);

void ans_renderer_destroy(ans_renderer* r);

/* ---- Scene ingest on the device. The reference loads a 3DGS PLY on the CPU
 * (aniso::loadPly, src/ply_loader.cpp:77-141: parse, activate, reorder f_rest) and then
 * uploads (gpu_render.cu:267-292). Here the host reads only the header, with the same rules and
 * messages (ply_loader.cpp:20-66, :86-110: binary_little_endian, float properties, columns by
 * name, SH degree from the f_rest count, length check); the vertex block goes to the device as
 * it is and one kernel activates it (exp, sigmoid, q/|q| in FP32, bit-identical to the
 * reference loader), drops the normals and writes the renderer's layout. Degree-4 files are
 * accepted and keep coefficients up to degree 3, the most either renderer evaluates
 * (sh.cpp:18-58). Malformed files: ANS_ERR_ARGUMENT with the reference loader's message. */

/* Header only, no GPU needed: the Gaussian count and SH degree of the file image data[bytes]. */
int ans_ply_inspect(const void* data, size_t bytes, int64_t* n, int32_t* sh_degree);

/* A renderer from a PLY file image in host memory (data[bytes], the whole file). */
int ans_renderer_create_ply(const void* data, size_t bytes, ans_renderer** out);

/* A renderer from a PLY file (loadPly(path), ply_loader.h:27): the vertex block is read in
 * chunks into pinned memory, each chunk's copy to the device overlapping the read of the next. */
int ans_renderer_load_ply(const char* path, ans_renderer** out);

/* The handle's scene copied back to host arrays in ans_renderer_create's layout (any array may
 * be NULL: call with all NULL to get n and sh_degree first). */
int ans_renderer_scene(ans_renderer* r, int64_t* n, int32_t* sh_degree, float* pos_xyz, float* scale_xyz,
                       float* rot_wxyz, float* opacity, float* sh);

/* Renders one frame into rgb_out (width * height * 3 bytes, RGB8, rows top to bottom), like
 * GpuRenderer::render (gpu_render.cu:299-375). options and stats may be NULL. An empty view
 * (nothing visible) gives a black image and status 0. */
int ans_render(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* rgb_out,
               size_t rgb_capacity, ans_stats* stats);

/* Same, writing the unclamped colour as float [height][width][3] (the image before the [0, 1]
 * clamp and quantization). Either output may be NULL. */
int ans_render_float(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* rgb_out,
                     float* rgbf_out, size_t pixels_capacity, ans_stats* stats);

/* Asynchronous variant for pipelined callers (the reference's `aniso path` loop renders one
 * view after another, main.cpp:234-244): enqueues the frame and returns without waiting.
 * Consecutive frames rotate over three internal streams with their own frame buffers, so a
 * frame's preprocess and sorts overlap the previous frames' blends; blends run in enqueue order,
 * so frames finish (and write their outputs) in the order they were enqueued. dev_rgb
 * (width*height*3 bytes) and dev_rgbf (width*height*3 floats) are DEVICE pointers; either may
 * be NULL. Call ans_render_wait before reading them. */
int ans_render_async(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* dev_rgb,
                     float* dev_rgbf);

/* Waits for every frame enqueued on the handle and fills stats: counts of the last frame, stage
 * times averaged over the frames enqueued since the previous wait (the newest 64 of them).
 * Returns ANS_ERR_CAPACITY if any of them had more pairs than the pair buffers held; the buffers
 * have then been grown and the caller renders those frames again. */
int ans_render_wait(ans_renderer* r, ans_stats* stats);

/* Intermediates of the last frame rendered on this handle, copied to host buffers (any may be
 * NULL). Sizes: splat_bits [n][10] uint32 = float bits of {u, v, conic A, B, C, opacity, depth,
 * r, g, b} (zeros for culled Gaussians); rect [n][4] int32 tile rect {x0, y0, x1, y1}
 * (x1 < x0 when culled); keys [pairs] uint64 = tile << 32 | depth bits; values [pairs] uint32
 * Gaussian index; ranges [tiles][2] uint32 [first, last) into the sorted pairs ([0, 0) if empty);
 * the keys and ranges of gpu_render.cu:157 and :164-177. */
int ans_frame_counts(ans_renderer* r, int64_t* n, int64_t* visible, int64_t* pairs, int64_t* tiles);
int ans_frame_dump(ans_renderer* r, uint32_t* splat_bits, int32_t* rect, uint64_t* keys, uint32_t* values,
                   uint32_t* ranges);

/* ---- Batched multi-view forward (no counterpart in the reference; its nearest caller renders
 * a camera path one GpuRenderer::render at a time, main.cpp:234-244). Renders n_views
 * (1..256) cameras of the handle's scene, each of its own size, with one sequence of batched
 * launches replayed as a captured HIP graph: every view's Gaussians are projected by one
 * preprocess launch, each sort runs all views' segments at once, and the blend takes every
 * view's tiles in one launch, longest tile list first. Each view's image equals
 * ans_render_async of that camera with the same options. dev_rgb[v] is a DEVICE pointer to
 * cams[v].width * height * 3 bytes. The batch runs on `stream` (a hipStream_t; NULL: the
 * default stream), after the caller's earlier work there, and the call waits for it; it grows
 * its pair arena and renders again if a view's pairs exceeded it. stats: NULL or n_views entries
 * (visible, tile_pairs and pair_capacity per view; blend_ms = the whole batch's GPU time;
 * project_ms = sort_ms = 0). */
int ans_render_views(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                     uint8_t* const* dev_rgb, void* stream, ans_stats* stats);

/* ans_render_views without the wait, for callers that render batch after batch (a training or
 * serving loop): the batch is enqueued on `stream` and the call returns. Two batches may be in
 * flight; a third call first waits for the oldest. A batch waits on the device for the batch in
 * flight (they share the pair arena); a batch that needs a larger arena waits on the host for
 * every batch in flight first. Pairs above the arena are not rendered again here: see
 * ans_views_wait. The async counterpart of ans_render_views as ans_render_async is of ans_render
 * (the reference renders one view per synchronous GpuRenderer::render, gpu_render.h:30). */
int ans_render_views_async(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                           uint8_t* const* dev_rgb, void* stream);

/* Waits for every batch of ans_render_views_async. Returns ANS_ERR_CAPACITY if a view of any of
 * them had more pairs than the arena held: the arena has then been grown and the caller renders
 * those batches again. stats: NULL or one entry per view of the last batch enqueued by
 * ans_render_views_async (as ans_render_views fills them, retries = 0). */
int ans_views_wait(ans_renderer* r, ans_stats* stats);

/* Tries every tile shape (8x8, 16x16, 32x16) and LDS batch depth (64, 128, 256) on single frames
 * of this camera (sh_degree, splat_scale and flags from base), each rendered into the DEVICE
 * buffer dev_rgb and waited for, and returns the fastest options in *best and their device time
 * (stamp-timed stages, the least of `repeats` frames after warm-up frames) in *best_ms. No
 * counterpart in the reference, whose tile size is fixed (gpu_render.cu:15). */
int ans_render_autotune(ans_renderer* r, const ans_camera* cam, const ans_options* base, uint8_t* dev_rgb,
                        int32_t repeats, ans_options* best, double* best_ms);

/* Tries every tile shape (8x8, 16x16, 32x16) and LDS batch depth (64, 128, 256) on this batch
 * (sh_degree, splat_scale and flags from base) and returns the fastest options in *best and
 * their batch time (the least of `repeats` timed replays after a warm-up) in *best_ms. */
int ans_views_autotune(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* base,
                       uint8_t* const* dev_rgb, void* stream, int32_t repeats, ans_options* best, double* best_ms);

/* Per-view lists of the last ans_render_views call, for parity checks against single-view
 * renders: counts, then values [pairs] (Gaussian index per sorted pair) and ranges [tiles][2]. */
int ans_views_frame_counts(ans_renderer* r, int32_t view, int64_t* visible, int64_t* pairs, int64_t* tiles);
int ans_views_frame_dump(ans_renderer* r, int32_t view, uint32_t* values, uint32_t* ranges);

/* ---- Camera-path streaming. Replaces the frame loop of the reference's `aniso path` verb
 * (src/main.cpp:234-244: one GpuRenderer::render per pose, each RGB8 frame fwrite to stdout).
 * Renders n_frames cameras in chunks of `chunk` views (1..256) with the batched launch above and
 * hands every frame, in order, to `sink` as RGB8 rows top to bottom (width * height * 3 bytes in
 * pinned host memory, valid until sink returns). Two chunks are in flight: while the host runs
 * sink on chunk k, chunk k + 1 renders and the copy of chunk k to the host overlaps it.
 * splat_scales: NULL (options->splat_scale for every frame) or n_frames positive scales, the
 * reference's optional tenth path column (main.cpp:219, :237). Each frame equals ans_render of
 * its camera with its splat scale. sink returns 0 to continue; nonzero stops the path and the
 * call returns ANS_ERR_ARGUMENT after its GPU work has drained. stats may be NULL. */
typedef int (*ans_frame_sink)(void* user, int64_t frame, const uint8_t* rgb, size_t bytes);
typedef struct ans_path_stats {
    int64_t frames;    /* frames handed to sink */
    int64_t chunks;    /* batched launches, re-renders after pair-arena growth included */
    double gpu_ms;     /* GPU time of the batched launches */
    double total_ms;   /* wall time of the call */
} ans_path_stats;
int ans_render_path(ans_renderer* r, int64_t n_frames, const ans_camera* cams, const float* splat_scales,
                    const ans_options* options, int32_t chunk, ans_frame_sink sink, void* user,
                    ans_path_stats* stats);

/* ---- Differentiable path (no counterpart in the reference, whose training calls a
 * third-party rasterizer at pipeline/train_gsplat.py:96-102; this is what a
 * torch.autograd.Function binds in its place). All pointers below are DEVICE pointers to
 * contiguous float32 arrays; `stream` is the hipStream_t to run on (NULL: the default stream), so
 * the kernels order with the caller's producers and consumers of those arrays. */

/* Raw parameters as an optimizer holds them: means [n][3], log_scales [n][3], quats [n][4]
 * (w, x, y, z; normalized inside), opacity_logits [n], sh [n][(sh_degree+1)^2][3]; quats and
 * sh 16-byte aligned. activated = 1: log_scales holds standard deviations and opacity_logits
 * opacities in [0, 1], the arguments the reference's training script passes to its rasterizer
 * after exp and sigmoid (pipeline/train_gsplat.py:98-101); the gradients are then w.r.t. those. */
typedef struct ans_params {
    int64_t n;
    int32_t sh_degree;
    const float* means;
    const float* log_scales;
    const float* quats;
    const float* opacity_logits;
    const float* sh;
    int32_t activated;
} ans_params;

/* Gradient outputs, same shapes as ans_params; overwritten (not accumulated). quats must be
 * 16-byte aligned. */
typedef struct ans_grads {
    float* means;
    float* log_scales;
    float* quats;
    float* opacity_logits;
    float* sh;
} ans_grads;

typedef struct ans_diff ans_diff;

int ans_diff_create(ans_diff** out);
void ans_diff_destroy(ans_diff* d);

/* Forward: renders the unclamped colour image [height][width][3] into `image` and keeps the
 * frame's state (splat records, tile lists) for the backward pass under *frame. The call waits
 * only for the pair count (gpu_render.cu:329 reads the same total back); if it exceeds the pair
 * buffers they grow before any pair is written. The rest of the frame is enqueued on `stream`
 * and not waited for: work the caller enqueues after the call sees the finished image. A non-NULL
 * stats waits for the frame. */
int ans_diff_forward(ans_diff* d, const ans_params* params, const ans_camera* cam, const ans_options* options,
                     float* image, void* stream, int64_t* frame, ans_stats* stats);

/* Backward for a frame from ans_diff_forward: gradients of L w.r.t. every raw parameter, given
 * grad_image = dL/dimage [height][width][3]. The frame stays valid until released. */
int ans_diff_backward(ans_diff* d, int64_t frame, const ans_params* params, const float* grad_image,
                      const ans_grads* grads, void* stream);

/* ans_diff_forward's image without allocating anything, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops). No counterpart in the reference. Every
 * per-frame buffer lives in one fixed 64 MiB array in device memory that the library holds on
 * each device for the life of the process: at most 524,288 Gaussians, and the pairs that fit in
 * what the Gaussians and tiles leave (stats->pair_capacity); a frame with more pairs returns
 * ANS_ERR_CAPACITY and leaves `image` unwritten. Runs on `stream` and waits for it. Calls on one
 * device share that array: calls that overlap in time on one device race. Keeps no frame for a
 * backward pass. */
int ans_diff_forward_noalloc(const ans_params* params, const ans_camera* cam, const ans_options* options,
                             float* image, void* stream, ans_stats* stats);

/* Releases a frame of ans_diff_forward or ans_diff_views_forward. */
int ans_diff_release(ans_diff* d, int64_t frame);

/* After ans_diff_backward on a frame: per-Gaussian gradients w.r.t. its splat record, copied to
 * host [n][9] = dL/d{u, v, conic A, B, C, opacity, r, g, b} (u, v: the screen-space mean
 * gradient that densification heuristics read). Valid until the next backward on the handle. */
int ans_diff_splat_grads(ans_diff* d, int64_t frame, float* out);

/* What a densification strategy reads next to the image (gsplat's `info`, consumed at
 * pipeline/train_gsplat.py:177, :185), as DEVICE arrays on `stream`: per-Gaussian splat radius in
 * pixels, 0 if culled (int32 [n]; any time after ans_diff_forward), and after ans_diff_backward on
 * the frame, the screen-space mean gradient dL/d(u, v) (float32 [n][2], 8-byte aligned; valid
 * until the next backward on the handle). */
int ans_diff_radii(ans_diff* d, int64_t frame, int32_t* radii, void* stream);
int ans_diff_screen_grads(ans_diff* d, int64_t frame, float* grad_uv, void* stream);

/* Per-tile lists of a frame (sizes from ans_diff_frame_counts), for gradient checks against a
 * CPU reference evaluated on the same lists: values [pairs], ranges [tiles][2]. */
int ans_diff_frame_counts(ans_diff* d, int64_t frame, int64_t* visible, int64_t* pairs, int64_t* tiles);
int ans_diff_frame_dump(ans_diff* d, int64_t frame, uint32_t* values, uint32_t* ranges);

/* ---- Batched differentiable frames (no counterpart in the reference, whose training loop
 * renders and differentiates one view per step through a third-party rasterizer,
 * pipeline/train_gsplat.py:170-190 with the render at :96-102). n_views (1..256) cameras of
 * uneven sizes render one parameter set in one sequence of batched launches replayed as a
 * captured HIP graph (the batched render's sequence, ans_render_views, on raw parameters); the
 * backward is one more graph: pair slots per view, one blend-adjoint launch over every view's
 * tiles, and one per-Gaussian launch that adds the views' gradients in view order (no atomics).
 * Each view's image equals ans_diff_forward of that camera; the gradients equal the sum over
 * views, in view order, of the single-view ans_diff_backward gradients. */

/* Forward: images[v] is a DEVICE float [height][width][3] for cams[v]. Runs on `stream` after
 * the caller's earlier work there; the call waits only until every view's pair count is known
 * (the pair arena grows and the batch is rendered again if a view's pairs exceeded it), and the
 * rest of the batch runs on: work the caller enqueues on `stream` afterwards sees the finished
 * images. stats: NULL or n_views entries (visible, tile_pairs, pair_capacity per view; blend_ms =
 * the batch's GPU time; a non-NULL stats waits for the batch). The frame (*frame) stays valid
 * until ans_diff_release. */
int ans_diff_views_forward(ans_diff* d, const ans_params* params, int32_t n_views, const ans_camera* cams,
                           const ans_options* options, float* const* images, void* stream, int64_t* frame,
                           ans_stats* stats);

/* Backward of a batched frame: grad_images[v] = dL/dimage of view v (DEVICE, as the images);
 * gradients of every raw parameter (overwritten) summed over the views; screen_grads: NULL or a
 * DEVICE float [n_views][n][2] (8-byte aligned) receiving dL/d(u, v) per view, what a
 * densification strategy reads (pipeline/train_gsplat.py:177, :185). Enqueued on `stream`, not
 * waited for. Backward passes of one handle must not overlap (they share scratch memory). */
int ans_diff_views_backward(ans_diff* d, int64_t frame, const ans_params* params, const float* const* grad_images,
                            const ans_grads* grads, float* screen_grads, void* stream);

/* Per view, 1 - final transmittance of every pixel (the alpha image, DEVICE float [height][width]);
 * it carries no gradient. */
int ans_diff_views_alpha(ans_diff* d, int64_t frame, float* const* alphas, void* stream);

/* Per (view, Gaussian) splat radius in pixels, 0 if culled in that view: DEVICE int32 [n_views][n]. */
int ans_diff_views_radii(ans_diff* d, int64_t frame, int32_t* radii, void* stream);

/* Per-view lists of a batched frame, for parity checks: counts, then values [pairs] and ranges
 * [tiles][2]. */
int ans_diff_views_counts(ans_diff* d, int64_t frame, int32_t view, int64_t* visible, int64_t* pairs,
                          int64_t* tiles);
int ans_diff_views_dump(ans_diff* d, int64_t frame, int32_t view, uint32_t* values, uint32_t* ranges);

/* The raw parameters (device arrays, as for ans_diff_forward) as a 3DGS PLY file image in host
 * memory, the training pipeline's writer (pipeline/train_gsplat.py:109-128): header, then per
 * Gaussian x y z, nx ny nz = 0, f_dc_0..2, f_rest channel-major, opacity logit, log-scales,
 * quaternion / max(|q|, 1e-12). out NULL: *written = the file size, nothing else done. Otherwise
 * out must hold capacity >= that size; the rows are interleaved by a kernel on `stream` and the
 * call waits for them. */
int ans_ply_write(const ans_params* params, void* out, size_t capacity, size_t* written, void* stream);

/* ---- COLMAP cameras and PNG frames (the reference CLI's files around the renderer) ---------- */

/* COLMAP's binary sparse model, sparse_dir/cameras.bin and sparse_dir/images.bin, as one camera per
 * registered image sorted by image name: aniso::loadColmapCameras (include/aniso/camera.h:44,
 * src/camera.cpp:59-140). SIMPLE_PINHOLE and PINHOLE only; other models, truncated files, unknown
 * models and images naming a missing camera fail with ANS_ERR_ARGUMENT and the reference's
 * message. *count = the number of images; cameras NULL: nothing else is written. Otherwise
 * cameras[capacity] receives them (ANS_ERR_CAPACITY if capacity < *count) and, unless names is
 * NULL, names[i * name_capacity] each NUL-terminated name. */
int ans_colmap_load(const char* sparse_dir, ans_camera* cameras, char* names, size_t name_capacity, int capacity,
                    int* count);

/* PNG file image of an RGB8 frame (rows top to bottom) in host memory, byte for byte the file
 * aniso::writePng writes (include/aniso/image.h:22, src/image.cpp:41-85): stored deflate blocks,
 * no compression. out NULL: *written = the file size only. */
int ans_png_encode(const uint8_t* rgb, int width, int height, void* out, size_t capacity, size_t* written);

/* aniso::writePng: the same bytes written to path. */
int ans_png_write(const char* path, const uint8_t* rgb, int width, int height);

/* Sizes for ans_png_encode_device: the PNG file and the device workspace it needs. */
int ans_png_device_sizes(int width, int height, size_t* file_bytes, size_t* workspace_bytes);

/* The PNG file image of a DEVICE RGB8 frame into DEVICE memory out[capacity], the same bytes as
 * ans_png_encode, computed by kernels enqueued on `stream` (scanline copy into the stored blocks,
 * Adler-32 and CRC-32 as parallel reductions); returns without waiting. workspace: device memory
 * of the size ans_png_device_sizes gives, used until the kernels end. The first call on a device
 * uploads a 5 KB constant table (a synchronous copy); later calls only enqueue kernels.
 * No counterpart in the reference, which encodes on the host after copying the frame back. */
int ans_png_encode_device(const void* rgb, int width, int height, void* out, size_t capacity, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Host activation used by ans_renderer_create_raw, exposed for parity tests. */
void ans_activate(int64_t n, const float* log_scale_xyz, const float* opacity_logit, const float* quat_wxyz,
                  float* scale_xyz, float* opacity, float* rot_wxyz);

#ifdef __cplusplus
}
#endif

#endif /* ANISO_AMD_H */
