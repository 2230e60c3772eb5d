"""Seeded synthetic scenes, cameras and workloads for the training step.

The reference ships no data (no fixtures, no scenes), so every workload here is
generated from a seed. Random numbers come from a counter-based SplitMix64
stream (one 64-bit hash per element index), turned into uniforms in (0, 1] and
then normals by Box-Muller; being counter-based, it vectorises in numpy and is
identical on any machine.

Scene layout is the reference's ``Gaussian3D`` record (math/mod.rs:27-46):
59 float32 per Gaussian = position[3], opacity_logit, rotation[4] (w, x, y, z),
log_scale[3], sh_coeffs[48] channel-major (``sh[ch*16 + k]``, sh.rs:93-104).
Cameras follow ``Camera`` (math/mod.rs:130-145): pinhole, row-major OpenCV
camera-to-world (+z forward, +y down).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

G_FLOATS = 59
GRAD_FLOATS = 60

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based random stream: element i of stream s under seed k is
    splitmix64(splitmix64(k) ^ (s << 40) ^ i)."""

    def __init__(self, seed: int):
        self.key = int(_splitmix64(np.array([seed], dtype=np.uint64))[0])
        self.next_stream = 0

    def _uniform(self, count: int) -> np.ndarray:
        s = self.next_stream
        self.next_stream += 1
        idx = np.arange(count, dtype=np.uint64)
        base = np.uint64(self.key ^ ((s << 40) & 0xFFFFFFFFFFFFFFFF))
        z = _splitmix64(idx ^ base)
        # 53 random bits -> (0, 1]
        return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)

    def normal(self, shape, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
        n = int(np.prod(shape))
        u1 = self._uniform(n)
        u2 = self._uniform(n)
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
        return (mean + std * z).reshape(shape)


def make_scene(n: int, seed: int, log_scale_mean: float = -4.0, log_scale_std: float = 0.5,
               sh_rest_std: float = 0.05, clusters: int = 0, cluster_sigma: float = 0.2) -> np.ndarray:
    """Return an (n, 59) float32 array of Gaussian3D records."""
    r = Stream(seed)
    g = np.zeros((n, G_FLOATS), dtype=np.float64)
    if clusters > 0:
        centres = r.normal((clusters, 3))
        member = r.normal((n, 3), std=cluster_sigma)
        g[:, 0:3] = centres[np.arange(n) % clusters] + member
    else:
        g[:, 0:3] = r.normal((n, 3))
    g[:, 3] = r.normal((n,))
    q = r.normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    g[:, 4:8] = q
    g[:, 8:11] = r.normal((n, 3), mean=log_scale_mean, std=log_scale_std)
    sh = np.zeros((n, 48))
    if sh_rest_std > 0:
        sh[:] = r.normal((n, 48), std=sh_rest_std)
    else:
        r.next_stream += 2
    dc = r.normal((n, 3), std=0.3)
    sh[:, 0] = dc[:, 0]
    sh[:, 16] = dc[:, 1]
    sh[:, 32] = dc[:, 2]
    g[:, 11:59] = sh
    return np.ascontiguousarray(g.astype(np.float32))


@dataclasses.dataclass
class Camera:
    """Pinhole camera (math/mod.rs:130-145); ``c2w`` row-major 4x4."""
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    c2w: np.ndarray  # (4, 4) float32

    def position(self) -> np.ndarray:
        return self.c2w[:3, 3].copy()


def look_at_origin(direction: np.ndarray, radius: float, width: int, height: int, focal: float) -> Camera:
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    centre = radius * d
    f = -d
    up = np.array([0.0, 1.0, 0.0])
    if abs(float(f @ up)) > 0.99:
        right = np.array([1.0, 0.0, 0.0])
    else:
        right = np.cross(f, up)
        right /= np.linalg.norm(right)
    down = np.cross(f, right)
    c2w = np.eye(4)
    c2w[:3, 0] = right
    c2w[:3, 1] = down
    c2w[:3, 2] = f
    c2w[:3, 3] = centre
    return Camera(width, height, float(focal), float(focal), width / 2.0, height / 2.0,
                  c2w.astype(np.float32))


def identity_camera(w: int, h: int) -> Camera:
    """tests/training_reduces_loss.rs:10-27 and backward.rs:828-846."""
    return Camera(w, h, float(w), float(w), w / 2.0, h / 2.0, np.eye(4, dtype=np.float32))


def make_cameras(seed: int, sizes, radius: float = 4.0, focal_scale: float | None = None,
                 focal: float | None = None):
    r = Stream(seed)
    dirs = r.normal((len(sizes), 3))
    cams = []
    for (w, h), d in zip(sizes, dirs):
        f = focal if focal is not None else focal_scale * w
        cams.append(look_at_origin(d, radius, w, h, f))
    return cams


@dataclasses.dataclass
class Workload:
    name: str
    n: int
    sizes: list
    focal: float | None
    focal_scale: float | None
    log_scale_mean: float = -4.0
    sh_rest_std: float = 0.05
    clusters: int = 0
    index: int = 0

    def scene(self, seed_offset: int = 1000) -> np.ndarray:
        return make_scene(self.n, seed_offset + self.index, log_scale_mean=self.log_scale_mean,
                          sh_rest_std=self.sh_rest_std, clusters=self.clusters)

    def target_scene(self) -> np.ndarray:
        return self.scene(2000)

    def cameras(self):
        return make_cameras(3000 + self.index, self.sizes, focal=self.focal, focal_scale=self.focal_scale)


_C2_SIZES = [(640, 480), (960, 540), (1280, 720)]

WORKLOADS = {
    "c0_10k_256": Workload("c0_10k_256", 10_000, [(256, 256)], 300.0, None, sh_rest_std=0.0, index=0),
    "c1_100k_800": Workload("c1_100k_800", 100_000, [(800, 800)], 1100.0, None, index=1),
    "c2_500k_8views": Workload("c2_500k_8views", 500_000, [_C2_SIZES[i % 3] for i in range(8)], None, 1.375,
                               index=2),
    "c3_1m_1080p_4views": Workload("c3_1m_1080p_4views", 1_000_000, [(1920, 1080)] * 4, None, 1.375, index=3),
    "c4_3m_clustered_4views": Workload("c4_3m_clustered_4views", 3_000_000, [(1920, 1080)] * 4, None, 1.375,
                                       log_scale_mean=-2.5, clusters=64, index=4),
}


def scaled(w: Workload, n: int | None = None, sizes=None) -> Workload:
    """A copy of a workload with fewer Gaussians and/or other view sizes (same
    distributions and seeds), for tests that must finish quickly."""
    out = dataclasses.replace(w)
    if n is not None:
        out.n = n
    if sizes is not None:
        out.sizes = list(sizes)
    return out


def crop_camera(cam: Camera, x0: int, y0: int, width: int, height: int) -> Camera:
    """The window [x0, x0 + width) x [y0, y0 + height) of a camera's image as
    a camera of its own: same pose and focal lengths, principal point moved
    by the window's origin."""
    return Camera(int(width), int(height), cam.fx, cam.fy, cam.cx - x0, cam.cy - y0, cam.c2w.copy())


def cluster_crops(w: Workload, size: int = 128):
    """For a clustered workload, one size x size window per view, centred on
    the projection of the cluster centre nearest to that view's camera among
    those projecting at least size / 2 inside its image: the densest part of
    the view (every tile there lists tens of thousands of Gaussians).
    Returns the cropped cameras."""
    assert w.clusters > 0
    centres = Stream(1000 + w.index).normal((w.clusters, 3))  # make_scene's first draw
    out = []
    for cam in w.cameras():
        c2w = cam.c2w.astype(np.float64)
        rot, t = c2w[:3, :3], c2w[:3, 3]
        pc = (centres - t) @ rot  # world -> camera: R^T (x - t)
        z = pc[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            sx = cam.fx * pc[:, 0] / z + cam.cx
            sy = cam.fy * pc[:, 1] / z + cam.cy
        m = size / 2
        ok = (z > 0.5) & (sx > m) & (sx < cam.width - m) & (sy > m) & (sy < cam.height - m)
        if not ok.any():
            raise ValueError("no cluster centre projects inside the view")
        k = np.where(ok)[0][np.argmin(z[ok])]
        x0 = int(round(sx[k])) - size // 2
        y0 = int(round(sy[k])) - size // 2
        out.append(crop_camera(cam, x0, y0, size, size))
    return out


def train_config(**over) -> dict:
    """TrainConfig defaults (pipeline/optimizer.rs:42-63)."""
    cfg = dict(lr_position=0.00016, lr_rotation=0.001, lr_scale=0.005, lr_opacity=0.05, lr_sh_dc=0.0025,
               lr_sh_rest=0.000125, beta1=0.9, beta2=0.999, eps=1e-8, densify_grad_threshold=0.0002,
               prune_opacity_thresh=0.005, densify_from_iter=500, densify_until_iter=15000,
               opacity_reset_iter=3000, max_gaussians=3_000_000, num_iterations=30000)
    cfg.update(over)
    return cfg


def benchmark_train_config() -> dict:
    """The workloads' learning rates: SH uses 5e-3 for the DC slots and 2.5e-4
    for the rest (the reference's 20x rule), the others as TrainConfig."""
    return train_config(lr_sh_dc=5e-3, lr_sh_rest=2.5e-4)
