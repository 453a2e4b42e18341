"""Scenes and cameras on the host: the synthetic scene recipe, the 3DGS PLY layout, and the
pinhole camera the renderer takes.

The PLY layout is the one the reference reads and writes (src/ply_loader.cpp:77-141,
pipeline/train_gsplat.py:109-128): binary little-endian float32, columns
``x y z nx ny nz f_dc_0..2 f_rest_0..(3K-4) opacity scale_0..2 rot_0..3`` with ``f_rest``
stored channel-major. Raw values are the optimizer's (log-scales, opacity logits, unnormalized
quaternions); activation happens in the renderer library's host code.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SH_C0 = 0.28209479177387814


@dataclasses.dataclass
class RawScene:
    """A scene in the raw 3DGS parameterization (what a PLY file and an optimizer hold).

    means [n,3], log_scales [n,3], quats [n,4] as (w, x, y, z), opacity_logits [n],
    sh [n, (degree+1)^2, 3] laid out [gaussian][coefficient][channel] (gaussian.h:33-39).
    """

    means: np.ndarray
    log_scales: np.ndarray
    quats: np.ndarray
    opacity_logits: np.ndarray
    sh: np.ndarray

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    @property
    def sh_degree(self) -> int:
        return int(round(math.sqrt(self.sh.shape[1]))) - 1

    def subset(self, idx) -> "RawScene":
        return RawScene(*(np.ascontiguousarray(a[idx]) for a in dataclasses.astuple(self)))


@dataclasses.dataclass
class Camera:
    """Pinhole camera in COLMAP's convention (camera.h:16-26): X_cam = R X_world + t, +x right,
    +y down the image, +z forward. Intrinsics in pixels, kept in double like the reference."""

    R: np.ndarray
    t: np.ndarray
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int

    def centre(self) -> np.ndarray:
        """World position of the camera centre, -R^T t (camera.cpp:19-23), in double."""
        return -(self.R.T @ self.t)

    def resized(self, width: int) -> "Camera":
        """The same camera at another width (camera.cpp:25-32)."""
        s = float(width) / self.width
        height = int(math.floor(self.height * s + 0.5)) if self.height * s >= 0 else int(round(self.height * s))
        return Camera(self.R.copy(), self.t.copy(), self.fx * s, self.fy * s, self.cx * s, self.cy * s, width, height)


def pinhole_camera(width: int, height: int, hfov_deg: float = 60.0) -> Camera:
    """Camera at the origin looking down +z with fx = fy = 0.5 W / tan(hfov / 2), principal
    point at the image centre: the single-image COLMAP model of the synthetic scenes."""
    f = 0.5 * width / math.tan(hfov_deg * 3.14159265358979 / 360.0)
    return Camera(np.eye(3), np.zeros(3), f, f, width / 2.0, height / 2.0, width, height)


def pose_camera(position, forward, up, width: int, height: int, hfov_deg: float) -> Camera:
    """A camera at `position` looking along `forward` with `up` roughly up, COLMAP convention
    (main.cpp:179-202): rows of R are right, down, forward, computed in float32 like the
    reference's Vec3 arithmetic."""
    f32 = np.float32

    def norm(v):
        v = np.asarray(v, dtype=f32)
        length = f32(math.sqrt(float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
        return (v / length).astype(f32)

    def cross(a, b):
        return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=f32)

    fwd = norm(forward)
    right = norm(cross(fwd, np.asarray(up, dtype=f32)))
    down = cross(fwd, right)
    rows = [right, down, fwd]
    pos = np.asarray(position, dtype=f32)
    R = np.array([[float(c) for c in r] for r in rows], dtype=np.float64)
    t = np.array([-float(r[0] * pos[0] + r[1] * pos[1] + r[2] * pos[2]) for r in rows], dtype=np.float64)
    fpx = 0.5 * width / math.tan(hfov_deg * 3.14159265358979 / 360.0)
    return Camera(R, t, fpx, fpx, width / 2.0, height / 2.0, width, height)


def synthetic_scene(n: int, width: int, height: int, seed: int = 0, clustered: bool = False,
                    n_clusters: int = 50) -> RawScene:
    """Synthetic SH3 scene. Draw order with numpy's default_rng(seed):

    z = U(2, 40); x = U(-1, 1) z tan 30deg; y = U(-1, 1) z tan 30deg H/W   (uniform in the frustum)
    [clustered: the first n/2 as above; 50 centres drawn the same way (z, x, y);
     k = integers(0, 50, n/2); points = centre[k] + N(0, 3^2)]
    log-scales N(-4.5, 0.8); quaternions N(0, 1) normalized; opacity logits N(0, 1.5);
    f_dc N(0, 0.5); f_rest N(0, 0.1).
    """
    rng = np.random.default_rng(seed)
    t = math.tan(math.radians(30))

    def frustum(m):
        z = rng.uniform(2, 40, m)
        x = rng.uniform(-1, 1, m) * z * t
        y = rng.uniform(-1, 1, m) * z * t * height / width
        return np.stack([x, y, z], 1)

    if clustered:
        half = n // 2
        first = frustum(half)
        centres = frustum(n_clusters)
        k = rng.integers(0, n_clusters, n - half)
        second = centres[k] + rng.normal(0, 3, (n - half, 3))
        means = np.concatenate([first, second], 0)
    else:
        means = frustum(n)
    log_scales = rng.normal(-4.5, 0.8, (n, 3))
    q = rng.normal(0, 1, (n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    logits = rng.normal(0, 1.5, n)
    f_dc = rng.normal(0, 0.5, (n, 3))
    f_rest = rng.normal(0, 0.1, (n, 45))
    # The file stores f_rest channel-major; in memory the layout is [coefficient][channel].
    sh = np.concatenate([f_dc[:, None, :], f_rest.reshape(n, 3, 15).transpose(0, 2, 1)], 1)
    f4 = np.float32
    return RawScene(means.astype(f4), log_scales.astype(f4), q.astype(f4), logits.astype(f4),
                    np.ascontiguousarray(sh.astype(f4)))


def _columns(sh_degree: int):
    rest = 3 * ((sh_degree + 1) ** 2 - 1)
    return (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + ["f_rest_%d" % i for i in range(rest)]
            + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def ply_bytes(scene: RawScene) -> bytes:
    """The scene as a 3DGS PLY file (header as pipeline/train_gsplat.py:124-128 writes it)."""
    n, k = scene.n, scene.sh.shape[1]
    f_rest = scene.sh[:, 1:, :].transpose(0, 2, 1).reshape(n, 3 * (k - 1))
    data = np.concatenate([scene.means, np.zeros((n, 3), np.float32), scene.sh[:, 0, :], f_rest,
                           scene.opacity_logits[:, None], scene.log_scales, scene.quats], 1).astype("<f4")
    header = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    header += "".join("property float %s\n" % c for c in _columns(scene.sh_degree)) + "end_header\n"
    return header.encode() + data.tobytes()


def write_ply(path, scene: RawScene) -> None:
    with open(path, "wb") as f:
        f.write(ply_bytes(scene))


class PlyError(ValueError):
    """Malformed or unsupported PLY input (mirrors aniso::PlyError, ply_loader.h:10-12)."""


def read_ply(path) -> RawScene:
    """Reads a binary little-endian 3DGS PLY (ply_loader.cpp:20-141 accepts the same files):
    all vertex properties float32; SH degree from the f_rest count."""
    with open(path, "rb") as f:
        if not f.readline().startswith(b"ply"):
            raise PlyError(f"{path}: not a PLY file")
        props, count, in_vertex, fmt = [], 0, False, None
        while True:
            line = f.readline()
            if not line:
                raise PlyError(f"{path}: header has no end_header")
            words = line.decode("ascii", "replace").strip().split()
            if not words:
                continue
            if words[0] == "format":
                fmt = words[1] if len(words) > 1 else ""
                if fmt != "binary_little_endian":
                    raise PlyError(f"{path}: only binary_little_endian PLY is supported, got {fmt}")
            elif words[0] == "element":
                in_vertex = words[1] == "vertex"
                if in_vertex:
                    count = int(words[2])
                elif int(words[2]) != 0:
                    raise PlyError(f"{path}: unexpected element '{words[1]}'")
            elif words[0] == "property" and in_vertex:
                if words[1] not in ("float", "float32"):
                    raise PlyError(f"{path}: property '{words[2]}' has type {words[1]}, expected float")
                props.append(words[2])
            elif words[0] == "end_header":
                if fmt is None:
                    raise PlyError(f"{path}: missing format line")
                break
        body = np.fromfile(f, dtype="<f4", count=count * len(props))
    if body.size != count * len(props):
        raise PlyError(f"{path}: file is shorter than the header promises ({count} Gaussians)")
    body = body.reshape(count, len(props))
    col = {p: i for i, p in enumerate(props)}

    def get(names):
        try:
            return np.ascontiguousarray(body[:, [col[c] for c in names]])
        except KeyError as e:
            raise PlyError(f"{path}: missing property '{e.args[0]}'") from None

    rest = 0
    while "f_rest_%d" % rest in col:
        rest += 1
    degree = next((d for d in range(5) if rest == 3 * ((d + 1) ** 2 - 1)), None)
    if degree is None:
        raise PlyError(f"{path}: {rest} f_rest properties do not match any SH degree")
    k = (degree + 1) ** 2
    dc = get(["f_dc_0", "f_dc_1", "f_dc_2"])
    f_rest = get(["f_rest_%d" % i for i in range(rest)]).reshape(count, 3, k - 1).transpose(0, 2, 1)
    sh = np.ascontiguousarray(np.concatenate([dc[:, None, :], f_rest], 1), dtype=np.float32)
    return RawScene(get(["x", "y", "z"]), get(["scale_0", "scale_1", "scale_2"]),
                    get(["rot_0", "rot_1", "rot_2", "rot_3"]), get(["opacity"])[:, 0].copy(), sh)


VIEW_SIZES = ((640, 360), (1280, 720), (1920, 1080))


def view_batch_cameras(count: int = 16, seed: int = 1, hfov_deg: float = 60.0):
    """`count` cameras of mixed sizes for the batched multi-view workload, drawn with numpy's
    default_rng(seed): size uniform over 640x360, 1280x720, 1920x1080; position uniform in the
    box [-1, 1]^3 around the origin; looking along +z turned by a yaw uniform in [-30, 30] degrees
    about the vertical axis, with image +y pointing down as in the synthetic scenes' camera."""
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(count):
        w, h = VIEW_SIZES[int(rng.integers(0, len(VIEW_SIZES)))]
        pos = rng.uniform(-1.0, 1.0, 3)
        yaw = math.radians(rng.uniform(-30.0, 30.0))
        fwd = (math.sin(yaw), 0.0, math.cos(yaw))
        cams.append(pose_camera(pos, fwd, (0.0, -1.0, 0.0), w, h, hfov_deg))
    return cams


def flythrough_cameras(count: int, width: int = 1920, height: int = 1080, hfov_deg: float = 70.0):
    """A smooth camera path of `count` frames through the synthetic scenes, the kind of path
    `aniso path` renders for a fly-through video (tools/flythrough.py: 1920x1080, 70 degrees):
    the camera dollies from (-0.4, 0, -0.4) to (0.4, 0.1, 0.4) while its yaw sweeps from -25 to
    25 degrees and back, eased at both ends."""
    cams = []
    for k in range(count):
        t = k / max(count - 1, 1)
        e = 0.5 - 0.5 * math.cos(math.pi * t)  # ease in and out
        pos = (-0.4 + 0.8 * e, 0.1 * e, -0.4 + 0.8 * e)
        yaw = math.radians(-25.0 * math.cos(2.0 * math.pi * t))
        fwd = (math.sin(yaw), 0.0, math.cos(yaw))
        cams.append(pose_camera(pos, fwd, (0.0, -1.0, 0.0), width, height, hfov_deg))
    return cams


# The benchmark workloads, by what they are.
WORKLOADS = {
    "synthetic-200k-640x360": dict(n=200_000, width=640, height=360, clustered=False),
    "synthetic-741883-1280x720": dict(n=741_883, width=1280, height=720, clustered=False),
    "clustered-2.2M-1280x720": dict(n=2_200_000, width=1280, height=720, clustered=True),
    "clustered-2.2M-1920x1080": dict(n=2_200_000, width=1920, height=1080, clustered=True),
    # 16 cameras of mixed sizes (view_batch_cameras); the scene fills the 1280x720 frustum.
    "batched-1M-16views": dict(n=1_000_000, width=1280, height=720, clustered=False, views=16),
}
