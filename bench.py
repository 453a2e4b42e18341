"""Benchmark: Gaussian-splat forward rasterization on one MI355X.

Workload: 741,883 synthetic SH3 Gaussians (FP32) at 1280x720, one pinhole camera, generated by
the recipe in aniso_amd.scenes (numpy default_rng(0)). A step is one forward frame: preprocess,
depth sort, pair expansion, pair sort, tile ranges and the blend, from the scene resident in
HBM to an RGB8 image in HBM. Warmup frames size the buffers; the timed frames are enqueued
back to back and bracketed by device synchronization.

Prints one JSON line (rank 0). `--gpus N` with N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame fwd and fwd+bwd @1280×720 (0.74M/2.2M Gaussians); views/s batched"
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="synthetic-741883-1280x720")
    ap.add_argument("--skip-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-extra", action="store_true", help="only the headline forward workload")
    ap.add_argument("--only", default="",
                    help="comma-separated extra legs to run (fwd_bwd, clustered, views, views_fwd_bwd, path, "
                         "ingest, png); default: all")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB device copy (the reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from aniso_amd import _lib, scenes
    from aniso_amd.renderer import GpuRenderer, RenderOptions, RenderStats, camera_c

    wl = scenes.WORKLOADS[args.workload]
    raw = scenes.synthetic_scene(wl["n"], wl["width"], wl["height"], clustered=wl["clustered"])
    cam = scenes.pinhole_camera(wl["width"], wl["height"])
    W, H = cam.width, cam.height
    print(f"# workload {args.workload}: {raw.n} Gaussians, {W}x{H}", file=sys.stderr, flush=True)

    L = _lib.load()
    r = GpuRenderer(raw)
    opts = RenderOptions()
    dev_img = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    c_cam = camera_c(cam)
    c_opt = opts.c()

    if args.calibrate_copy:
        copy_calibration()

    # Warmup: sizes every buffer, including those of the library's three asynchronous streams
    # (frames rotate over them), through the same entry points as the timed region.
    st = RenderStats()
    r.render(cam, st, opts)
    for _ in range(max(3, args.warmup)):
        _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev_img.data_ptr(), None))
    _lib.check(L.ans_render_wait(r._h, None))

    # Timed region: K frames enqueued back to back, device-resident input and output. The
    # library's per-frame clock stamps (taken by the frame's kernels at each stage boundary) give
    # the stage times averaged over the timed frames (the newest 64), read by the wait.
    torch.cuda.synchronize()
    _lib.check(L.ans_render_wait(r._h, None))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev_img.data_ptr(), None))
    cst = _lib.Stats_()
    rc = L.ans_render_wait(r._h, ctypes.byref(cst))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.check(rc)
    ms = (t1 - t0) * 1e3 / args.steps
    proj, sort_, blend = cst.project_ms, cst.sort_ms, cst.blend_ms
    P, V = cst.tile_pairs, cst.visible
    T = ((W + 15) // 16) * ((H + 15) // 16)

    # Algorithmic HBM bytes per frame (compulsory unique traffic).
    n = raw.n
    b_fwd = 248 * n + 64 * V + 84 * P + 8 * T + 3 * W * H
    b_blend = 40 * P + 3 * W * H
    overlapped_ms = {"project": proj, "sort": sort_, "blend": blend}

    # The same frame through the host-output entry (ans_render: synchronous, image copied to host
    # memory over PCIe), as the reference's GpuRenderer::render returns it.
    host_img = np.empty((H, W, 3), np.uint8)
    host_steps = max(5, args.steps // 5)
    r.render(cam, None, opts)
    t0 = time.perf_counter()
    for _ in range(host_steps):
        _lib.check(L.ans_render(r._h, c_cam, c_opt, host_img.ctypes.data_as(ctypes.c_void_p), host_img.size, None))
    host_ms = (time.perf_counter() - t0) * 1e3 / host_steps

    extra = {"host_output": {"ms_per_frame": round(host_ms, 4), "steps": host_steps,
                             "includes": "synchronous ans_render: frame, pair-count readback, RGB8 copy to host"},
             "single_frame": bench_single_frames(r, L, c_cam, c_opt, dev_img)}
    sf = extra["single_frame"]
    # Stage costs from frames rendered one at a time (the pipelined frames overlap, so their stage
    # spans add up to more than a frame).
    stage_ms = sf.pop("stages_ms")
    dominant = max(stage_ms, key=stage_ms.get)
    roofline = blend_roofline(args.workload, sf["warm_blend_median_ms"], b_blend)
    roofline["overlapped"] = {
        "kernel_ms": round(blend, 4),
        "note": "the same kernel's stamp-timed mean in the timed frames, which overlap on the library's three "
                "streams (the blend shares the chip with the next frames' sorts)"}
    legs = set(x for x in args.only.split(",") if x) if args.only else None
    want = lambda leg: not args.skip_extra and (legs is None or leg in legs)
    if want("fwd_bwd"):
        extra["fwd_bwd"] = bench_fwd_bwd(raw, cam, max(5, args.steps // 4), args.warmup)
    if want("clustered"):
        extra["fwd_clustered_2.2M"] = bench_forward("clustered-2.2M-1280x720", max(10, args.steps // 2), args.warmup)
        extra["fwd_clustered_2.2M_1080p"] = bench_forward("clustered-2.2M-1920x1080", max(10, args.steps // 2),
                                                         args.warmup)
    if want("views"):
        extra["batched_16_views"] = bench_views("batched-1M-16views", max(5, args.steps // 10), args.warmup)
    if want("views_fwd_bwd"):
        extra["batched_16_views_fwd_bwd"] = bench_views_fwd_bwd("batched-1M-16views", max(5, args.steps // 10),
                                                                 args.warmup)
    if want("path"):
        extra["path_stream"] = bench_path(raw, 128)
    if want("ingest"):
        extra["ply_ingest"] = bench_ingest(raw)
    png_frame = None
    if want("png"):
        extra["png_1080p"], png_frame = bench_png(r)

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(raw, cam, args.cpu_seconds, png_frame)

    if cpu and "ply_loader" in cpu and "ply_ingest" in extra:
        extra["ply_ingest"]["speedup_vs_reference_loader"] = round(
            cpu["ply_loader"]["value"] / extra["ply_ingest"]["load_ply_ms"], 2)
    if cpu and "png_writer" in cpu and "png_1080p" in extra:
        extra["png_1080p"]["speedup_vs_reference_writer"] = round(
            cpu["png_writer"]["value"] / extra["png_1080p"]["device_to_file_ms"], 2)

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(sf["warm_median_ms"] / 2.3, 4),
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "gaussians": n, "width": W, "height": H, "sh_degree": 3,
                   "tile": "16x16", "batch_depth": opts.batch_depth, "visible": V, "pairs": P},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": stage_ms,
        "stages_basis": "median stage times of frames rendered one at a time (device clock stamps)",
        "stages_ms_overlapped": {k: round(v, 4) for k, v in overlapped_ms.items()},
        "dominant_stage": dominant,
        "frame_hbm": {"algorithmic_bytes": b_fwd, "achieved_GBs": round(b_fwd / (ms * 1e-3) / 1e9, 1)},
        "kernels": kernel_table(args.workload),
        "vs_baseline_note": "single_frame.warm_median_ms (one frame rendered and waited for, the way the "
                            "reference's render() is timed, main.cpp:161-162) / 2.3 ms, the reference's "
                            "published RTX 5090 forward time for its 741,883-Gaussian SH3 scene at 1280 wide "
                            "(README.md:26-28); lower is better",
        **extra,
    }
    print(json.dumps(out), flush=True)


COPY_BYTES = 1 << 30


def copy_calibration():
    """Three launches of a plain elementwise copy (torch.mul by 1.0) reading COPY_BYTES and writing
    COPY_BYTES: the known-size kernel whose FETCH_SIZE / WRITE_SIZE counters scale the counters
    of the rasterizer's kernels to bytes."""
    import torch

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def measured_traffic(kernel, workload, field="traffic_bytes_per_launch"):
    """A per-launch counter figure of `kernel` on `workload` from the committed counter summary
    profiles/<kernel>_<workload>.md (its `<field>:` line, by default the HBM bytes), or None."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{field}:\s*(\d+)", open(path).read(), re.M)
    return int(m.group(1)) if m else None


# The vector ALU of MI355X without packed FP32 (which this build keeps off): 256 CUs x 4 SIMDs,
# a wave64 instruction occupying a 16-lane SIMD for 4 cycles, 2.4 GHz peak engine clock.
SIMDS = 1024
VALU_CYCLES_PER_INST = 4
CLOCK_HZ = 2.4e9


def bench_single_frames(r, L, c_cam, c_opt, dev_img, warm=25, cold=10):
    """Frames rendered one at a time, each waited for: the median of `warm` frames after one
    warm-up, and the median of `cold` frames each preceded by writing 1 GiB of scratch (more than
    the last-level cache, so the scene and the frame's buffers come from HBM). Wall time per frame
    and the frame's GPU time from the library's clock stamps."""
    import torch

    from aniso_amd import _lib

    stages = []

    def one():
        t0 = time.perf_counter()
        _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev_img.data_ptr(), None))
        st = _lib.Stats_()
        _lib.check(L.ans_render_wait(r._h, ctypes.byref(st)))
        wall = (time.perf_counter() - t0) * 1e3
        stages.append((st.project_ms, st.sort_ms, st.blend_ms))
        return wall, st.project_ms + st.sort_ms + st.blend_ms, st.blend_ms

    one()
    stages.clear()
    warm_runs = [one() for _ in range(warm)]
    warm_stages = list(stages)
    scratch = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
    cold_runs = []
    for k in range(cold):
        scratch.fill_(k & 0xff)
        torch.cuda.synchronize()
        cold_runs.append(one())
    del scratch
    med = lambda xs: round(statistics.median(xs), 4)
    return {"warm_median_ms": med([w for w, _, _ in warm_runs]),
            "warm_gpu_median_ms": med([g for _, g, _ in warm_runs]),
            "warm_blend_median_ms": med([b for _, _, b in warm_runs]),
            "cold_median_ms": med([w for w, _, _ in cold_runs]),
            "cold_gpu_median_ms": med([g for _, g, _ in cold_runs]),
            "warm_runs": warm, "cold_runs": cold,
            "cold_flush": "1 GiB device write before each cold frame (evicts the last-level cache)",
            "includes": "ans_render_async + ans_render_wait per frame; gpu = stamp-timed project + sort + blend",
            "stages_ms": {k: med([x[i] for x in warm_stages]) for i, k in enumerate(("project", "sort", "blend"))}}


def profile_figure(kernel, workload, field):
    """A figure of `kernel` on `workload` from the committed counter summary
    profiles/<kernel>_<workload>.md (its `<field>:` line), or None."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{field}:\s*([0-9.eE+-]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def blend_roofline(workload, kernel_ms, b_blend):
    """The blend's roofline. Its bound is the vector ALU (about 20 instructions per visited splat and
    8x8 pixel block; its HBM traffic is 12% of peak), so the primary figure is the VALU issue rate:
    wave instructions per launch (SQ_INSTS_VALU, committed profile) over the isolated kernel time,
    against 1,024 SIMDs x 2.4 GHz / 4 cycles per wave64 FP32 instruction. HBM is the secondary
    bound: algorithmic bytes (40 per pair + 3 per pixel) over the same time, against 8 TB/s, with
    the measured traffic (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy)."""
    peak = SIMDS * CLOCK_HZ / VALU_CYCLES_PER_INST / 1e9
    inst = profile_figure("k_blend", workload, "valu_instructions_per_launch")
    achieved = inst / (kernel_ms * 1e-3) / 1e9 if inst else None
    hbm = b_blend / (kernel_ms * 1e-3) / 1e9
    frags = profile_figure("k_blend", workload, "fragments_per_frame")
    return {
        "bound": "valu",
        "kernel": "k_blend",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": round(peak, 1),
        "unit": "G wave64 VALU instructions/s",
        "frac": round(achieved / peak, 4) if achieved else None,
        "traffic": measured_traffic("k_blend", workload),
        "kernel_ms": round(kernel_ms, 4),
        "timing": "device clock stamps at the blend's start and end, median of frames rendered one at a time "
                  "(the rocprof kernel-trace mean of the same launches is in the committed profile)",
        "valu_instructions_per_launch": int(inst) if inst else None,
        "occupancy_waves_per_simd": profile_figure("k_blend", workload, "mean_occupancy_waves_per_simd"),
        "lds_bank_conflict_cycles": profile_figure("k_blend", workload, "lds_bank_conflict_cycles"),
        "fragments_per_frame": int(frags) if frags else None,
        "fragments_per_s": round(frags / (kernel_ms * 1e-3), 1) if frags else None,
        "hbm": {"achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm / HBM_PEAK_GBS, 4), "algorithmic_bytes": b_blend},
    }


def kernel_table(workload):
    """Per-kernel HBM rates of a frame and a training step from the committed counter profile
    profiles/kernels_<workload>.md: measured bytes per launch over the rocprof mean duration."""
    import re

    path = os.path.join(ROOT, "profiles", f"kernels_{workload}.md")
    if not os.path.exists(path):
        return None
    out = {}
    for m in re.finditer(r"^kernel: (\S+) us=([0-9.]+) fetch=([0-9]+) write=([0-9]+)", open(path).read(), re.M):
        name, us, fetch, write = m.group(1), float(m.group(2)), int(m.group(3)), int(m.group(4))
        gbs = (fetch + write) / (us * 1e-6) / 1e9
        out[name] = {"us": us, "hbm_bytes": fetch + write, "GBs": round(gbs, 1),
                     "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def bench_forward(workload, steps, warmup):
    """Forward ms/frame of another synthetic workload, same protocol as the headline line, with
    frames rendered one at a time beside it and the per-kernel HBM rates of its committed profile."""
    import torch

    from aniso_amd import _lib, scenes
    from aniso_amd.renderer import GpuRenderer, RenderOptions, camera_c

    wl = scenes.WORKLOADS[workload]
    raw = scenes.synthetic_scene(wl["n"], wl["width"], wl["height"], clustered=wl["clustered"])
    cam = scenes.pinhole_camera(wl["width"], wl["height"])
    L = _lib.load()
    with GpuRenderer(raw) as r:
        dev = torch.empty(cam.width * cam.height * 3, dtype=torch.uint8, device="cuda")
        c_cam, c_opt = camera_c(cam), RenderOptions().c()
        for _ in range(max(3, warmup)):
            _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev.data_ptr(), None))
        _lib.check(L.ans_render_wait(r._h, None))
        t0 = time.perf_counter()
        for _ in range(steps):
            _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev.data_ptr(), None))
        cst = _lib.Stats_()
        _lib.check(L.ans_render_wait(r._h, ctypes.byref(cst)))
        ms = (time.perf_counter() - t0) * 1e3 / steps
        single = bench_single_frames(r, L, c_cam, c_opt, dev, warm=15, cold=5)
    return {"ms_per_frame": round(ms, 4), "gaussians": raw.n, "visible": cst.visible, "pairs": cst.tile_pairs,
            "width": cam.width, "height": cam.height, "steps": steps,
            "stages_ms_overlapped": {"project": round(cst.project_ms, 4), "sort": round(cst.sort_ms, 4),
                                     "blend": round(cst.blend_ms, 4)},
            "single_frame": single, "kernels": kernel_table(workload)}


def bench_views(workload, steps, warmup):
    """Views/s of the batched multi-view forward (one ans_render_views call per batch of 16
    mixed-size cameras, graph replay, autotuned tile shape and LDS batch depth) against the same
    16 cameras rendered by single-view launches, all into device images, on one scene. Two
    single-view baselines: one complete render per view, each waited for (the reference's path
    loop, main.cpp:234-244, calls render() once per view), and all 16 enqueued back to back with
    one wait at the end (pipelined)."""
    import torch

    from aniso_amd import _lib, scenes
    from aniso_amd.renderer import GpuRenderer, RenderOptions, camera_c

    wl = scenes.WORKLOADS[workload]
    raw = scenes.synthetic_scene(wl["n"], wl["width"], wl["height"], clustered=wl["clustered"])
    cams = scenes.view_batch_cameras(wl["views"])
    L = _lib.load()
    outs = [torch.empty(c.width * c.height * 3, dtype=torch.uint8, device="cuda") for c in cams]
    with GpuRenderer(raw) as r:
        # Single-view launches at the default options (16x16 tiles, depth 256).
        c_opt = RenderOptions().c()
        c_cams = [camera_c(c) for c in cams]

        def plain():
            for c, o in zip(c_cams, outs):
                _lib.check(L.ans_render_async(r._h, c, c_opt, o.data_ptr(), None))
            _lib.check(L.ans_render_wait(r._h, None))

        def per_view():
            for c, o in zip(c_cams, outs):
                _lib.check(L.ans_render_async(r._h, c, c_opt, o.data_ptr(), None))
                _lib.check(L.ans_render_wait(r._h, None))

        def timed(fn):
            for _ in range(max(1, warmup)):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / steps

        plain_ms = timed(plain)
        per_view_ms = timed(per_view)

        best, tuned_ms = r.autotune_views(cams, outs, repeats=3)
        stats = []
        for _ in range(max(1, warmup)):
            r.render_views(cams, outs, best, stats)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.render_views(cams, outs, best)
        torch.cuda.synchronize()
        batch_ms = (time.perf_counter() - t0) * 1e3 / steps
        # Batches enqueued back to back (ans_render_views_async, two in flight), one wait at the
        # end: the host's per-batch work overlaps the previous batch on the GPU.
        for _ in range(max(1, warmup)):
            r.render_views_async(cams, outs, best)
        assert r.views_wait(), "pair arena overflow after the warm-up"
        t0 = time.perf_counter()
        for _ in range(steps):
            r.render_views_async(cams, outs, best)
        assert r.views_wait(), "pair arena overflow"
        async_ms = (time.perf_counter() - t0) * 1e3 / steps
        r.render_views(cams, outs, RenderOptions(tile=best.tile, batch_depth=best.batch_depth, graph=False))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.render_views(cams, outs, RenderOptions(tile=best.tile, batch_depth=best.batch_depth, graph=False))
        torch.cuda.synchronize()
        direct_ms = (time.perf_counter() - t0) * 1e3 / steps
    k = len(cams)
    return {"views": k, "gaussians": raw.n, "sizes": [f"{c.width}x{c.height}" for c in cams],
            "pairs_total": int(sum(s.tile_pairs for s in stats)),
            "batched_views_per_s": round(k / (batch_ms * 1e-3), 1), "batched_ms_per_batch": round(batch_ms, 4),
            "batched_direct_launch_ms_per_batch": round(direct_ms, 4),
            "single_view_ms_per_batch": round(per_view_ms, 4),
            "single_view_views_per_s": round(k / (per_view_ms * 1e-3), 1),
            "speedup": round(per_view_ms / batch_ms, 3),
            "speedup_basis": "16 single-view renders, each waited for (the reference's per-view render())",
            "single_view_pipelined_ms_per_batch": round(plain_ms, 4),
            "speedup_vs_pipelined": round(plain_ms / batch_ms, 3),
            "batched_async_ms_per_batch": round(async_ms, 4),
            "batched_async_views_per_s": round(k / (async_ms * 1e-3), 1),
            "speedup_async": round(per_view_ms / async_ms, 3),
            "speedup_async_vs_pipelined": round(plain_ms / async_ms, 3),
            "async_basis": "batches enqueued back to back with ans_render_views_async, one ans_views_wait at the "
                           "end, against the same single-view baselines",
            "autotuned": {"tile": f"{best.tile[0]}x{best.tile[1]}", "batch_depth": best.batch_depth,
                          "gpu_ms": round(tuned_ms, 4)}, "steps": steps}


def bench_path(raw, frames):
    """Frames/s of camera-path streaming: a `frames`-frame fly-through at 1920x1080, 70 degrees
    (scenes.flythrough_cameras, the size and field of view of tools/flythrough.py) of the
    headline scene, every RGB8 frame delivered in order into host memory. ans_render_path
    (chunks of 16 views in batched launches, two chunks in flight, the copy of one overlapping the
    render of the next) against the reference's loop (src/main.cpp:234-244): one synchronous
    ans_render per frame returning the host image."""
    from aniso_amd import _lib, scenes
    from aniso_amd.renderer import GpuRenderer, RenderOptions, camera_c

    cams = scenes.flythrough_cameras(frames)
    L = _lib.load()
    with GpuRenderer(raw) as r:
        opts = RenderOptions()
        r.render_path(cams[:32], lambda i, rgb: None, opts, chunk=16)  # graphs captured, buffers sized
        st = r.render_path(cams, lambda i, rgb: None, opts, chunk=16)
        path_ms = st.total_ms
        img = np.empty((cams[0].height, cams[0].width, 3), np.uint8)
        c_opt = opts.c()
        c_cams = [camera_c(c) for c in cams]
        _lib.check(L.ans_render(r._h, c_cams[0], c_opt, img.ctypes.data_as(ctypes.c_void_p), img.size, None))
        t0 = time.perf_counter()
        for c in c_cams:
            _lib.check(L.ans_render(r._h, c, c_opt, img.ctypes.data_as(ctypes.c_void_p), img.size, None))
        loop_ms = (time.perf_counter() - t0) * 1e3
    return {"frames": frames, "size": f"{cams[0].width}x{cams[0].height}", "gaussians": raw.n, "chunk": 16,
            "path_frames_per_s": round(frames / (path_ms * 1e-3), 1), "path_ms_per_frame": round(path_ms / frames, 4),
            "path_gpu_ms_per_frame": round(st.gpu_ms / frames, 4),
            "per_frame_loop_frames_per_s": round(frames / (loop_ms * 1e-3), 1),
            "per_frame_loop_ms_per_frame": round(loop_ms / frames, 4),
            "speedup": round(loop_ms / path_ms, 3),
            "includes": "every frame's RGB8 image in host memory, in order (pinned, PCIe)"}


def bench_ingest(raw):
    """Scene ingest from a 3DGS PLY file of the headline scene (written to a temporary file, so it
    is in the page cache): ans_renderer_load_ply (header on the host, the vertex block streamed
    through pinned memory, activation and reorder on the device) and ans_renderer_create_ply
    (from the file's bytes in memory), best of 3 each. The reference loader's time on the same
    file is part of the CPU baseline (cpu_baseline.ply_loader)."""
    import tempfile

    from aniso_amd import scenes
    from aniso_amd.renderer import GpuRenderer

    fd, path = tempfile.mkstemp(suffix=".ply")
    os.close(fd)
    try:
        scenes.write_ply(path, raw)
        size = os.path.getsize(path)
        data = open(path, "rb").read()

        def best(fn, k=3):
            out = 1e30
            for _ in range(k):
                t0 = time.perf_counter()
                r = fn()
                out = min(out, (time.perf_counter() - t0) * 1e3)
                r.close()
            return out

        load_ms = best(lambda: GpuRenderer.from_ply(path))
        bytes_ms = best(lambda: GpuRenderer.from_ply(data))
    finally:
        os.remove(path)
    return {"gaussians": raw.n, "file_bytes": size, "load_ply_ms": round(load_ms, 2),
            "load_ply_GBs": round(size / (load_ms * 1e-3) / 1e9, 2), "create_ply_from_bytes_ms": round(bytes_ms, 2),
            "includes": "file read (page cache) or bytes in host memory -> renderer ready on the device"}


def bench_png(r):
    """PNG file images of a rendered 1920x1080 frame of the headline scene (16x16 tiles), in the
    reference writer's layout (image.cpp:41-85): encoded on the device from the frame in HBM
    (ans_png_encode_device: kernel time from HIP events, best of 20), then with the copy of the file
    image to pinned host memory, against the host encoder (ans_png_encode) from the frame in host
    memory. The reference writer's time on the same frame is part of the CPU baseline
    (cpu_baseline.png_writer). Returns (stats, the frame in host memory)."""
    import torch

    from aniso_amd import files, scenes

    cam = scenes.pinhole_camera(1920, 1080)
    frame = torch.empty(1080 * 1920 * 3, dtype=torch.uint8, device="cuda")
    r.render_views([cam], [frame])
    torch.cuda.synchronize()
    img = frame.view(1080, 1920, 3)
    host_frame = img.cpu().numpy()
    out = torch.empty(8 << 20, dtype=torch.uint8, device="cuda")
    work = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    pinned = torch.empty(8 << 20, dtype=torch.uint8, pin_memory=True)
    enc = files.png_device(img, out, work)
    size = enc.numel()
    torch.cuda.synchronize()
    assert enc.cpu().numpy().tobytes() == files.png_bytes(host_frame)
    import tempfile

    kernel, full, to_file = [], [], []
    s = torch.cuda.current_stream()
    fd, path = tempfile.mkstemp(suffix=".png")
    os.close(fd)
    view = pinned[:size].numpy()
    try:
        for _ in range(20):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            files.png_device(img, out, work)
            e1.record(s)
            torch.cuda.synchronize()
            kernel.append(e0.elapsed_time(e1))
            t0 = time.perf_counter()
            files.png_device(img, out, work)
            pinned[:size].copy_(out[:size], non_blocking=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            with open(path, "wb") as f:
                f.write(view)
            t2 = time.perf_counter()
            full.append((t1 - t0) * 1e3)
            to_file.append((t2 - t0) * 1e3)
    finally:
        os.remove(path)
    host = []
    for _ in range(5):
        t0 = time.perf_counter()
        files.png_bytes(host_frame)
        host.append((time.perf_counter() - t0) * 1e3)
    k = min(kernel)
    return ({"width": 1920, "height": 1080, "file_bytes": size, "device_encode_ms": round(k, 4),
             "device_encode_GBs": round(size / (k * 1e-3) / 1e9, 1),
             "device_to_host_file_ms": round(min(full), 4), "device_to_file_ms": round(min(to_file), 4),
             "host_encode_ms": round(min(host), 3),
             "includes": "device: frame in HBM -> PNG file image in HBM (device_to_host_file_ms: and copied to pinned "
                         "host memory; device_to_file_ms: and written to a file in the page cache, as the reference "
                         "writer does); host: frame in host memory -> file image in host memory"},
            host_frame)


def bench_fwd_bwd(raw, cam, steps, warmup):
    """One differentiable training step on the headline scene: forward (unclamped float image), L1
    against a target rendered after perturbing means by N(0, 0.01^2) and SH DC by N(0, 0.05^2),
    backward to every raw parameter. Timed over `steps` steps between device synchronizations."""
    import torch

    from aniso_amd import scenes
    from aniso_amd.autograd import rasterize

    dev = "cuda"
    t = lambda a, g=False: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev, requires_grad=g)
    params = [t(a, True) for a in (raw.means, raw.log_scales, raw.quats, raw.opacity_logits, raw.sh)]
    rng = np.random.default_rng(0)
    tm = raw.means + rng.normal(0, 0.01, raw.means.shape).astype(np.float32)
    tsh = raw.sh.copy()
    tsh[:, 0, :] += rng.normal(0, 0.05, tsh[:, 0, :].shape).astype(np.float32)
    with torch.no_grad():
        target = rasterize(t(tm), t(raw.log_scales), t(raw.quats), t(raw.opacity_logits), t(tsh), cam)

    def step():
        for p in params:
            p.grad = None
        img = rasterize(*params, cam)
        loss = (img - target).abs().mean()
        loss.backward()
        return loss

    for _ in range(max(1, warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"ms_per_step": round(ms, 4), "steps": steps, "loss": "L1 mean vs perturbed-scene target",
            "includes": "forward + loss + backward to means, log-scales, quats, opacity logits, SH"}


def bench_views_fwd_bwd(workload, steps, warmup):
    """The batched training step (1.0M Gaussians, 16 cameras of uneven sizes): one batched
    differentiable forward of all 16 views (rasterize_views: one graph-captured launch sequence),
    the mean L1 loss against targets rendered after perturbing means by N(0, 0.01^2) and SH DC by
    N(0, 0.05^2), and one batched backward (one graph: blend adjoint over every view's tiles,
    per-Gaussian adjoint summing the views), against the same step as 16 single-view forward +
    backward passes (rasterize per view, gradients accumulated by autograd)."""
    import torch

    from aniso_amd import scenes
    from aniso_amd.autograd import rasterize, rasterize_views

    wl = scenes.WORKLOADS[workload]
    raw = scenes.synthetic_scene(wl["n"], wl["width"], wl["height"], clustered=wl["clustered"])
    cams = scenes.view_batch_cameras(wl["views"])
    dev = "cuda"
    t = lambda a, g=False: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev, requires_grad=g)
    params = [t(a, True) for a in (raw.means, raw.log_scales, raw.quats, raw.opacity_logits, raw.sh)]
    rng = np.random.default_rng(0)
    tm = raw.means + rng.normal(0, 0.01, raw.means.shape).astype(np.float32)
    tsh = raw.sh.copy()
    tsh[:, 0, :] += rng.normal(0, 0.05, tsh[:, 0, :].shape).astype(np.float32)
    with torch.no_grad():
        tflat = rasterize_views(t(tm), t(raw.log_scales), t(raw.quats), t(raw.opacity_logits), t(tsh), cams,
                                flat=True)
    V = len(cams)
    targets, o = [], 0
    for c in cams:
        targets.append(tflat[o:o + c.height * c.width * 3].view(c.height, c.width, 3))
        o += c.height * c.width * 3
    # The batched loss over the flat images: mean over views of each view's L1 mean, as one dot
    # product with the weights (1 / (V * 3 H W) for every value of a view).
    weight = torch.cat([torch.full((c.height * c.width * 3,), 1.0 / (V * 3 * c.height * c.width), device=dev)
                        for c in cams])
    stats = []

    def batched():
        for p in params:
            p.grad = None
        flat = rasterize_views(*params, cams, stats=stats if not stats else None, flat=True)
        loss = torch.dot((flat - tflat).abs(), weight)
        loss.backward()

    def per_view():
        for p in params:
            p.grad = None
        for cam, tg in zip(cams, targets):
            loss = (rasterize(*params, cam) - tg).abs().mean() / V
            loss.backward()

    def timed(fn):
        for _ in range(max(1, warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    per_view_ms = timed(per_view)
    batch_ms = timed(batched)
    return {"views": V, "gaussians": raw.n, "sizes": [f"{c.width}x{c.height}" for c in cams],
            "pairs_total": int(sum(s.tile_pairs for s in stats)),
            "batched_ms_per_step": round(batch_ms, 4), "batched_views_per_s": round(V / (batch_ms * 1e-3), 1),
            "per_view_ms_per_step": round(per_view_ms, 4),
            "per_view_views_per_s": round(V / (per_view_ms * 1e-3), 1),
            "speedup": round(per_view_ms / batch_ms, 3),
            "speedup_basis": "16 single-view differentiable forward + backward passes (rasterize per view)",
            "loss": "mean over views of L1 mean vs perturbed-scene targets", "steps": steps,
            "includes": "forward of all views + loss + backward to means, log-scales, quats, opacity logits, SH"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(raw, cam, seconds, png_frame=None):
    """The reference CPU renderer, compiled from its own sources (oracle/_ref), on the same
    scene and camera, all host threads (std::thread::hardware_concurrency, render.cpp:108).
    Falls back to the oracle's restatement in reference mode when the compiled library is
    absent. Median ms/frame of the frames that fit in `seconds` (at least 2) after a warm-up."""
    import oracle

    act = oracle.activate(raw)
    if oracle.ref_available():
        kind = "reference"
        fn = lambda: oracle.ref_render(act, cam)[0]
    else:
        kind = "port"
        fn = lambda: oracle.render(act, cam, oracle.REFERENCE, tile=(8, 8))
    fn()
    times = []
    t_end = time.perf_counter() + seconds
    while len(times) < 2 or (time.perf_counter() < t_end and len(times) < 50):
        t = time.perf_counter()
        fn()
        times.append((time.perf_counter() - t) * 1e3)
    out = {"value": round(statistics.median(times), 2), "unit": "ms/frame", "cores": os.cpu_count(),
           "cpu": cpu_model(), "kind": kind,
           "sample": f"{len(times)} frames of the same workload after 1 warm-up, median"}
    if oracle.ref_available():
        # The reference's PLY loader (aniso::loadPly, one thread) on the scene's file, beside the
        # device ingest leg.
        import tempfile

        from aniso_amd import scenes

        fd, path = tempfile.mkstemp(suffix=".ply")
        os.close(fd)
        try:
            scenes.write_ply(path, raw)
            ms = min(oracle.ref_load_ply(path)[1] for _ in range(2))
        finally:
            os.remove(path)
        out["ply_loader"] = {"value": round(ms, 2), "unit": "ms/file", "cores": 1, "kind": "reference",
                             "sample": f"aniso::loadPly of the {raw.n}-Gaussian SH3 PLY, best of 2"}
        if png_frame is not None:
            # The reference's PNG writer (aniso::writePng, one thread) on the frame of the png leg,
            # into a temporary file (page cache).
            fd, path = tempfile.mkstemp(suffix=".png")
            os.close(fd)
            try:
                ms = min(oracle.ref_write_png(path, png_frame) for _ in range(3))
            finally:
                os.remove(path)
            out["png_writer"] = {"value": round(ms, 2), "unit": "ms/file", "cores": 1, "kind": "reference",
                                 "sample": "aniso::writePng of the png leg's 1920x1080 frame, best of 3"}
    return out


if __name__ == "__main__":
    main()
