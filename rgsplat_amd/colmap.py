"""COLMAP sparse-model readers (pipeline/sfm.rs:132-350) and the SfM output
loader (pipeline/mod.rs:270-296), on the host: they parse files and hand the
points to gaussians_from_sfm (rgs_gaussians_from_sfm / rgs_scene_from_sfm)
and the posed cameras to the training step.

Binary layouts (little-endian), as the reference reads them:
  points3D.bin  u64 count; per point u64 id, 3 f64 xyz, 3 u8 rgb, f64 error,
                u64 track length, then track length x 8 bytes (skipped)
  cameras.bin   u64 count; per camera u32 id, i32 model, u64 width,
                u64 height, then the model's f64 parameters
  images.bin    u64 count; per image u32 id, 4 f64 qvec (w, x, y, z), 3 f64
                tvec, u32 camera id, NUL-terminated name, u64 point count,
                then point count x 24 bytes (skipped)
A file that ends early raises ValueError (the reference's read_exact fails).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from .synthetic import Camera

f32 = np.float32


class _Reader:
    def __init__(self, data: bytes):
        self.d = memoryview(data)
        self.o = 0

    def take(self, fmt: str):
        n = struct.calcsize(fmt)
        if self.o + n > len(self.d):
            raise ValueError("unexpected end of file")
        v = struct.unpack_from(fmt, self.d, self.o)
        self.o += n
        return v

    def skip(self, n: int):
        if self.o + n > len(self.d):
            raise ValueError("unexpected end of file")
        self.o += n


@dataclass
class SfMPoints:
    """SfMPoint (sfm.rs:132-140) as arrays."""
    ids: np.ndarray       # (n,) u64
    xyz: np.ndarray       # (n, 3) f64
    rgb: np.ndarray       # (n, 3) u8
    error: np.ndarray     # (n,) f64
    track_length: np.ndarray  # (n,) u64


def parse_points3d_bin(data: bytes) -> SfMPoints:
    """parse_points3d_bin (sfm.rs:147-185)."""
    r = _Reader(data)
    (n,) = r.take("<Q")
    ids, xyz, rgb, err, tl = [], [], [], [], []
    for _ in range(n):
        pid, x, y, z, cr, cg, cb, e, t = r.take("<Q3d3BdQ")
        r.skip(t * 8)
        ids.append(pid)
        xyz.append((x, y, z))
        rgb.append((cr, cg, cb))
        err.append(e)
        tl.append(t)
    return SfMPoints(np.array(ids, np.uint64), np.array(xyz, np.float64).reshape(-1, 3),
                     np.array(rgb, np.uint8).reshape(-1, 3), np.array(err, np.float64), np.array(tl, np.uint64))


def colmap_camera_params(model_id: int) -> int:
    """sfm.rs:342-351."""
    return {0: 3, 1: 4, 2: 4, 3: 5, 4: 8}.get(model_id, 4)


def parse_cameras_bin(data: bytes, width: int, height: int) -> dict[int, Camera]:
    """parse_cameras_bin (sfm.rs:188-231): camera id -> intrinsics with an
    identity pose. PINHOLE and OPENCV give fx, fy, cx, cy; the other models
    f, cx, cy; the camera's own size unless it is zero."""
    r = _Reader(data)
    (n,) = r.take("<Q")
    out = {}
    for _ in range(n):
        cam_id, model, w, h = r.take("<IiQQ")
        params = r.take(f"<{colmap_camera_params(model)}d")
        w, h = w & 0xFFFFFFFF, h & 0xFFFFFFFF  # `as u32`
        cw, ch = (w, h) if w > 0 and h > 0 else (width, height)
        if model in (1, 4) and len(params) >= 4:
            fx, fy, cx, cy = (float(f32(p)) for p in params[:4])
        elif len(params) >= 3:
            fx, fy, cx, cy = float(f32(params[0])), float(f32(params[0])), float(f32(params[1])), float(f32(params[2]))
        else:
            fx, fy, cx, cy = float(f32(cw)), float(f32(ch)), float(f32(cw) / f32(2)), float(f32(ch) / f32(2))
        out[cam_id] = Camera(int(cw), int(ch), fx, fy, cx, cy, np.eye(4, dtype=np.float32))
    return out


@dataclass
class ColmapImage:
    """ColmapImage (sfm.rs:238-245)."""
    image_id: int
    qvec: tuple
    tvec: tuple
    camera_id: int
    name: str


def parse_images_bin(data: bytes) -> list[ColmapImage]:
    """parse_images_bin (sfm.rs:252-297)."""
    r = _Reader(data)
    (n,) = r.take("<Q")
    out = []
    for _ in range(n):
        image_id, qw, qx, qy, qz, tx, ty, tz, cam_id = r.take("<I4d3dI")
        name = bytearray()
        while True:
            (b,) = r.take("<B")
            if b == 0:
                break
            name.append(b)
        (npts,) = r.take("<Q")
        r.skip(npts * 24)
        out.append(ColmapImage(image_id, (qw, qx, qy, qz), (tx, ty, tz), cam_id,
                               name.decode("utf-8", errors="replace")))
    return out


def _rotation_f32(q):
    """quaternion::to_rotation_matrix (math/quaternion.rs:9-37) in f32."""
    q = [f32(v) for v in q]
    n = np.sqrt(f32(f32(f32(q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]))
    w, x, y, z = (f32(1), f32(0), f32(0), f32(0)) if n < f32(1e-12) else (q[0] / n, q[1] / n, q[2] / n, q[3] / n)
    one, two = f32(1), f32(2)
    return np.array([[one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y)],
                     [two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x)],
                     [two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)]], f32)


def build_cameras(images: list[ColmapImage], intrinsics: dict[int, Camera], image_dir: str):
    """build_cameras (sfm.rs:303-340): c2w = [R^T | -R^T t] in f32; images
    of unknown cameras are skipped. Returns (cameras, image paths)."""
    cams, paths = [], []
    for img in images:
        intr = intrinsics.get(img.camera_id)
        if intr is None:
            continue
        rt = _rotation_f32(img.qvec).T
        t = [f32(v) for v in img.tvec]
        c2w = np.zeros((4, 4), f32)
        for row in range(3):
            c2w[row, :3] = rt[row]
            acc = rt[row, 0] * t[0]
            acc = f32(acc + rt[row, 1] * t[1])
            acc = f32(acc + rt[row, 2] * t[2])
            c2w[row, 3] = -acc
        c2w[3] = [0, 0, 0, 1]
        cams.append(Camera(intr.width, intr.height, intr.fx, intr.fy, intr.cx, intr.cy, c2w))
        paths.append(os.path.join(image_dir, img.name))
    return cams, paths


def load_sfm_output(workspace: str, image_dir: str, width: int = 1920, height: int = 1080):
    """load_sfm_output (pipeline/mod.rs:270-296) without the Gaussians: a
    model directory under workspace/sparse (the first in name order; the
    reference takes the first its directory listing yields), its points (to
    pass to gaussians_from_sfm) and its posed cameras with their image
    paths."""
    sparse = os.path.join(workspace, "sparse")
    dirs = sorted(d for d in os.listdir(sparse) if os.path.isdir(os.path.join(sparse, d)))
    if not dirs:
        raise FileNotFoundError(f"no COLMAP model found in {sparse}")
    model = os.path.join(sparse, dirs[0])

    def read(name):
        with open(os.path.join(model, name), "rb") as fh:
            return fh.read()

    pts = parse_points3d_bin(read("points3D.bin"))
    intr = parse_cameras_bin(read("cameras.bin"), width, height)
    cams, paths = build_cameras(parse_images_bin(read("images.bin")), intr, image_dir)
    return pts, cams, paths
