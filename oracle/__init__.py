"""ctypes wrapper over the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product package (rgsplat_amd)
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None

c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("c2w", ctypes.c_float * 16), ("distortion", ctypes.c_float * 3)]


class TrainCfg(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in ("lr_position", "lr_rotation", "lr_scale", "lr_opacity", "lr_sh_dc",
                                              "lr_sh_rest", "beta1", "beta2", "eps", "densify_grad_threshold",
                                              "prune_opacity_thresh")] + \
               [(k, ctypes.c_uint32) for k in ("densify_from_iter", "densify_until_iter", "opacity_reset_iter")] + \
               [(k, ctypes.c_uint64) for k in ("max_gaussians", "num_iterations")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(
                os.path.join(HERE, "oracle.cpp")):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, sz, i32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
    L.orc_run_f32.restype = vp
    L.orc_run_f32.argtypes = [vp, sz, ctypes.POINTER(Camera), vp, vp, ctypes.c_float, i32, ctypes.POINTER(i32)]
    L.orc_run_f64.restype = vp
    L.orc_run_f64.argtypes = [vp, sz, ctypes.POINTER(Camera), vp, vp, ctypes.c_double, i32, ctypes.POINTER(i32)]
    L.orc_free.argtypes = [vp, i32]
    for f in ("orc_nsplats", "orc_npairs"):
        getattr(L, f).restype = sz
        getattr(L, f).argtypes = [vp, i32]
    L.orc_pairs_fwd.restype = u64
    L.orc_pairs_fwd.argtypes = [vp, i32]
    L.orc_get_splats.argtypes = [vp, i32, vp, vp, vp]
    L.orc_get_tiles.argtypes = [vp, i32, vp, vp]
    L.orc_get_image.argtypes = [vp, i32, vp, vp, vp]
    L.orc_get_loss.restype = ctypes.c_double
    L.orc_get_loss.argtypes = [vp, i32]
    L.orc_get_loss_exact.restype = ctypes.c_double
    L.orc_get_loss_exact.argtypes = [vp, i32]
    L.orc_get_dl.argtypes = [vp, i32, vp]
    L.orc_get_grads.argtypes = [vp, i32, vp]
    L.orc_backward_with_dl_f32.argtypes = [vp, vp, sz, vp, vp, vp]
    L.orc_rasterize_rgba.restype = i32
    L.orc_rasterize_rgba.argtypes = [vp, sz, ctypes.POINTER(Camera), vp, vp]
    L.orc_loss_grad_f32_exact.restype = ctypes.c_double
    L.orc_loss_grad_f32_exact.argtypes = [vp, vp, sz, sz, ctypes.c_float, vp, ctypes.POINTER(ctypes.c_double)]
    L.orc_loss_grad_f32.restype = ctypes.c_double
    L.orc_loss_grad_f32.argtypes = [vp, vp, sz, sz, ctypes.c_float, vp]
    L.orc_loss_grad_f64.restype = ctypes.c_double
    L.orc_loss_grad_f64.argtypes = [vp, vp, sz, sz, ctypes.c_double, vp]
    L.orc_ssim_channel.restype = ctypes.c_double
    L.orc_ssim_channel.argtypes = [vp, vp, sz, sz, vp]
    L.orc_gaussian_kernel.argtypes = [vp]
    L.orc_blur.argtypes = [vp, sz, sz, vp]
    L.orc_sigmoid.restype = ctypes.c_float
    L.orc_sigmoid.argtypes = [ctypes.c_float]
    L.orc_logit.restype = ctypes.c_float
    L.orc_logit.argtypes = [ctypes.c_float]
    for f in ("orc_sh_basis", "orc_sh_basis_grad", "orc_q_normalize", "orc_mat_from_quat",
              "orc_to_rotation_matrix", "orc_rotation_jacobian", "orc_covariance_3d"):
        getattr(L, f).argtypes = [vp, vp]
    L.orc_eval_sh_deg3.argtypes = [vp, vp, vp]
    L.orc_build_cov3d.argtypes = [vp, vp, vp]
    L.orc_backprop_cov3d.argtypes = [vp, vp, vp, vp, vp]
    L.orc_project_covariance.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_float, vp]
    L.orc_aabb.restype = i32
    L.orc_aabb.argtypes = [ctypes.c_float, ctypes.c_float, vp, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, vp]
    L.orc_linear_to_srgb.restype = ctypes.c_float
    L.orc_linear_to_srgb.argtypes = [ctypes.c_float]
    L.orc_srgb_to_linear.restype = ctypes.c_float
    L.orc_srgb_to_linear.argtypes = [ctypes.c_float]
    L.orc_powi.restype = ctypes.c_float
    L.orc_powi.argtypes = [ctypes.c_float, i32]
    L.orc_view_matrix.restype = i32
    L.orc_view_matrix.argtypes = [ctypes.POINTER(Camera), vp]
    L.orc_adam_step.argtypes = [vp, vp, sz, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(TrainCfg)]
    L.orc_reset_opacities.argtypes = [vp, sz]
    L.orc_scene_extent.restype = ctypes.c_float
    L.orc_scene_extent.argtypes = [vp, sz]
    L.orc_densify_and_prune.restype = i32
    L.orc_densify_and_prune.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), sz, vp, vp, vp, vp,
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(TrainCfg), ctypes.c_float]
    L.orc_densify_and_prune_split.restype = i32
    L.orc_densify_and_prune_split.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), sz, vp, vp, vp, vp,
                                              ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(TrainCfg),
                                              ctypes.c_float, ctypes.c_float]
    L.orc_set_threads.argtypes = [i32]
    L.orc_expf_batch.argtypes = [vp, vp, sz]
    L.orc_gaussians_from_sfm.restype = i32
    L.orc_gaussians_from_sfm.argtypes = [vp, vp, sz, vp, vp]
    for f in ("orc_preview_project",):
        getattr(L, f).restype = i32
        getattr(L, f).argtypes = [vp, sz, ctypes.POINTER(Camera), vp]
    L.orc_resize_rgba8.restype = i32
    L.orc_resize_rgba8.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint32]
    L.orc_preview_render.restype = i32
    L.orc_preview_render.argtypes = [vp, sz, ctypes.POINTER(Camera), vp, vp]
    L.orc_rgba8_to_linear.argtypes = [vp, sz, vp]
    L.orc_to_u8_srgb.argtypes = [vp, sz, vp]
    L.orc_save_ply.restype = i32
    L.orc_save_ply.argtypes = [vp, sz, ctypes.c_char_p]
    L.orc_load_ply.restype = i32
    L.orc_load_ply.argtypes = [ctypes.c_char_p, vp, sz, ctypes.POINTER(ctypes.c_size_t)]
    L.orc_save_splat.restype = i32
    L.orc_save_splat.argtypes = [vp, sz, ctypes.c_char_p]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def camera(cam) -> Camera:
    c = Camera()
    c.width, c.height = int(cam.width), int(cam.height)
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    for i, v in enumerate(np.asarray(cam.c2w, dtype=np.float32).reshape(16)):
        c.c2w[i] = float(v)
    return c


def train_cfg(d: dict) -> TrainCfg:
    c = TrainCfg()
    for k, _ in TrainCfg._fields_:
        setattr(c, k, d[k])
    return c


def set_threads(n: int):
    lib().orc_set_threads(int(n))


class Render:
    """Result of backward.rs:107-149 (render_and_backward) or, without a
    target, :152-166 (render_color), with its intermediates."""

    def __init__(self, g, cam, target=None, bg=(0.0, 0.0, 0.0), lam=0.2, f64=False, backward=True):
        L = lib()
        dt = np.float64 if f64 else np.float32
        self.f64 = int(f64)
        self.dtype = dt
        self.n = len(g)
        self.w, self.h = int(cam.width), int(cam.height)
        self._g = np.ascontiguousarray(g, dtype=dt)
        self._t = None if target is None else np.ascontiguousarray(target, dtype=dt)
        self._bg = np.ascontiguousarray(bg, dtype=dt)
        self._cam = camera(cam)
        err = ctypes.c_int(0)
        fn = L.orc_run_f64 if f64 else L.orc_run_f32
        h = fn(_p(self._g), self.n, ctypes.byref(self._cam), None if self._t is None else _p(self._t),
               _p(self._bg), lam, int(backward and target is not None), ctypes.byref(err))
        if not h:
            raise ValueError(f"oracle render failed (code {err.value})")
        self.h_ = h
        self.has_backward = backward and target is not None

    def __del__(self):
        if getattr(self, "h_", None):
            lib().orc_free(self.h_, self.f64)
            self.h_ = None

    @property
    def n_splats(self):
        return lib().orc_nsplats(self.h_, self.f64)

    @property
    def n_pairs(self):
        return lib().orc_npairs(self.h_, self.f64)

    @property
    def pairs_fwd(self):
        return lib().orc_pairs_fwd(self.h_, self.f64)

    def splats(self):
        ns = self.n_splats
        gi = np.zeros(ns, np.uint64)
        f = np.zeros((ns, 22), self.dtype)
        rects = np.zeros((ns, 4), np.int32)
        lib().orc_get_splats(self.h_, self.f64, _p(gi), _p(f), _p(rects))
        names = ["sx", "sy", "conic", "cov2d", "color", "raw_color", "opacity", "depth", "pcam"]
        out = {"gi": gi, "rect": rects, "sx": f[:, 0], "sy": f[:, 1], "conic": f[:, 2:5], "cov2d": f[:, 5:8],
               "color": f[:, 8:11], "raw_color": f[:, 11:14], "opacity": f[:, 14], "depth": f[:, 15],
               "pcam": f[:, 19:22]}
        del names
        return out

    def tiles(self):
        ntx = (self.w + 15) // 16
        nty = (self.h + 15) // 16
        off = np.zeros(ntx * nty + 1, np.uint64)
        gis = np.zeros(max(self.n_pairs, 1), np.uint32)
        lib().orc_get_tiles(self.h_, self.f64, _p(off), _p(gis))
        return off, gis[: self.n_pairs]

    def image(self):
        rgb = np.zeros((self.h, self.w, 3), self.dtype)
        t = np.zeros((self.h, self.w), self.dtype)
        last = np.zeros((self.h, self.w), np.int32)
        lib().orc_get_image(self.h_, self.f64, _p(rgb), _p(t), _p(last))
        return rgb, t, last

    @property
    def loss(self):
        return lib().orc_get_loss(self.h_, self.f64)

    @property
    def loss_exact(self):
        """The loss from the same per-pixel terms added in double (the
        reference adds them serially in the working precision)."""
        return lib().orc_get_loss_exact(self.h_, self.f64)

    def dl(self):
        d = np.zeros((self.h, self.w, 3), self.dtype)
        lib().orc_get_dl(self.h_, self.f64, _p(d))
        return d

    def grads(self):
        gr = np.zeros((self.n, 60), self.dtype)
        lib().orc_get_grads(self.h_, self.f64, _p(gr))
        return gr

    def backward_with_dl(self, dl):
        """The reference backward (backward_pass + geometry_backprop) of this
        forward with the given dL/dpixel (h, w, 3) instead of its own loss's
        (f32 handles only)."""
        assert not self.f64
        dl = np.ascontiguousarray(dl, np.float32).reshape(-1)
        assert dl.size == self.w * self.h * 3
        gr = np.zeros((self.n, 60), np.float32)
        lib().orc_backward_with_dl_f32(self.h_, _p(self._g), self.n, _p(self._bg), _p(dl), _p(gr))
        return gr


def render_and_backward(g, cam, target, bg=(0.0, 0.0, 0.0), lam=0.2, f64=False):
    r = Render(g, cam, target, bg, lam, f64=f64)
    return r.loss, r.grads()


def render_color(g, cam, bg=(0.0, 0.0, 0.0), f64=False):
    r = Render(g, cam, None, bg, f64=f64, backward=False)
    return r.image()[0]


def rasterize_rgba(g, cam, bg=(0.0, 0.0, 0.0)):
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros((cam.height, cam.width, 4), np.float32)
    bgv = np.ascontiguousarray(bg, np.float32)
    c = camera(cam)
    rc = lib().orc_rasterize_rgba(_p(g), len(g), ctypes.byref(c), _p(bgv), _p(out))
    if rc:
        raise ValueError("singular camera")
    return out


def loss_grad(rgb, target_rgba, w, h, lam, f64=False):
    dt = np.float64 if f64 else np.float32
    rgb = np.ascontiguousarray(rgb, dt)
    tg = np.ascontiguousarray(target_rgba, dt)
    dl = np.zeros(w * h * 3, dt)
    fn = lib().orc_loss_grad_f64 if f64 else lib().orc_loss_grad_f32
    loss = fn(_p(rgb), _p(tg), w, h, lam, _p(dl))
    return loss, dl


def loss_grad_exact(rgb, target_rgba, w, h, lam):
    """loss_grad in f32 -> (loss, the same per-pixel terms added in double,
    dL/drgb)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    tg = np.ascontiguousarray(target_rgba, np.float32)
    dl = np.zeros(w * h * 3, np.float32)
    ex = ctypes.c_double()
    loss = lib().orc_loss_grad_f32_exact(_p(rgb), _p(tg), w, h, lam, _p(dl), ctypes.byref(ex))
    return loss, ex.value, dl


def view_matrix(cam) -> np.ndarray:
    out = np.zeros(16, np.float32)
    c = camera(cam)
    if lib().orc_view_matrix(ctypes.byref(c), _p(out)):
        raise ValueError("singular c2w")
    return out.reshape(4, 4)


class AdamState:
    """AdamState (optimizer.rs:74-106) with a fixed row capacity."""

    def __init__(self, n, cap=None):
        cap = cap or n
        self.m1 = np.zeros((cap, 59), np.float32)
        self.m2 = np.zeros((cap, 59), np.float32)
        self.accum = np.zeros(cap, np.float32)
        self.denom = np.zeros(cap, np.uint32)
        self.rows = n
        self.t = 0


def adam_step(g, grads, st: AdamState, cfg: dict):
    assert g.dtype == np.float32 and g.flags.c_contiguous
    grads = np.ascontiguousarray(grads, np.float32)
    t = ctypes.c_uint64(st.t)
    c = train_cfg(cfg)
    lib().orc_adam_step(_p(g), _p(grads), len(g), _p(st.m1), _p(st.m2), _p(st.accum), _p(st.denom),
                        ctypes.byref(t), ctypes.byref(c))
    st.t = t.value


def reset_opacities(g):
    lib().orc_reset_opacities(_p(g), len(g))


def scene_extent(g):
    g = np.ascontiguousarray(g, np.float32)
    return lib().orc_scene_extent(_p(g), len(g))


def densify_and_prune(g_buf, n, st: AdamState, cfg: dict, extent: float, split_div: float = 0.0):
    """g_buf: (cap, 59) float32 buffer whose first n rows are live. Returns new n.
    split_div > 0 replaces the reference's split scale (s0 / sqrt 2 on axis
    0) by every scale / split_div."""
    nn = ctypes.c_size_t(n)
    rows = ctypes.c_size_t(st.rows)
    c = train_cfg(cfg)
    rc = lib().orc_densify_and_prune_split(_p(g_buf), ctypes.byref(nn), len(g_buf), _p(st.m1), _p(st.m2),
                                           _p(st.accum), _p(st.denom), ctypes.byref(rows), ctypes.byref(c),
                                           float(extent), float(split_div))
    if rc:
        raise ValueError("capacity exceeded")
    st.rows = rows.value
    return nn.value


def expf_batch(x):
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.zeros_like(x)
    lib().orc_expf_batch(_p(x), _p(out), len(x))
    return out


def f32vec(lst):
    return np.ascontiguousarray(lst, np.float32)


def call3(name, *arrays, out_shape):
    out = np.zeros(out_shape, np.float32)
    getattr(lib(), name)(*[_p(f32vec(a)) for a in arrays], _p(out))
    return out


# ---- model files (io/ply.rs) -------------------------------------------------
def save_ply(g, path):
    """save_ply (io/ply.rs:35-102)."""
    g = np.ascontiguousarray(g, np.float32).reshape(-1, 59)
    if lib().orc_save_ply(_p(g), len(g), str(path).encode()):
        raise OSError(f"oracle save_ply failed: {path}")


def load_ply(path, capacity=None):
    """load_ply (io/ply.rs:116-247): (n, 59) Gaussian3D records. Raises
    ValueError for what the reference rejects (bad header or short body)."""
    n = ctypes.c_size_t(0)
    # first pass sizes the buffer
    rc = lib().orc_load_ply(str(path).encode(), None, 0, ctypes.byref(n))
    if rc not in (0, 4):
        raise ValueError(f"oracle load_ply failed ({rc}): {path}")
    cap = n.value if capacity is None else capacity
    out = np.zeros((max(cap, 1), 59), np.float32)
    rc = lib().orc_load_ply(str(path).encode(), _p(out), cap, ctypes.byref(n))
    if rc:
        raise ValueError(f"oracle load_ply failed ({rc}): {path}")
    return out[: n.value]


def save_splat(g, path):
    """save_splat (io/ply.rs:258-303)."""
    g = np.ascontiguousarray(g, np.float32).reshape(-1, 59)
    if lib().orc_save_splat(_p(g), len(g), str(path).encode()):
        raise OSError(f"oracle save_splat failed: {path}")


# ---- colour conversions (io/images.rs:63-82, cpu/rasterizer.rs:38-49) -------
def rgba8_to_linear(img):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros(img.shape, np.float32)
    lib().orc_rgba8_to_linear(_p(img), img.size // 4, _p(out))
    return out


def to_u8_srgb(rgba):
    rgba = np.ascontiguousarray(rgba, np.float32)
    out = np.zeros(rgba.shape, np.uint8)
    lib().orc_to_u8_srgb(_p(rgba), rgba.size // 4, _p(out))
    return out


# ---- SfM initialisation (pipeline/sfm.rs:365-455) ---------------------------
def gaussians_from_sfm(xyz, rgb):
    """(Gaussian3D records (n, 59), per-point KNN scales (n,))."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    rgb = np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3)
    n = len(xyz)
    out = np.zeros((max(n, 1), 59), np.float32)
    sc = np.zeros(max(n, 1), np.float32)
    if lib().orc_gaussians_from_sfm(_p(xyz), _p(rgb), n, _p(out), _p(sc)):
        raise ValueError("NaN coordinate (the reference panics)")
    return out[:n], sc[:n]


def preview_project(g, cam):
    """GpuRenderer::project (gpu/compute.rs:138-207): (n, 10) SplatInfo2D
    records (sx, sy, conic a, b, c, colour r, g, b, alpha, depth)."""
    g = np.ascontiguousarray(g, np.float32).reshape(-1, 59)
    out = np.zeros((max(len(g), 1), 10), np.float32)
    c = camera(cam)
    if lib().orc_preview_project(_p(g), len(g), ctypes.byref(c), _p(out)):
        raise ValueError("singular camera")
    return out[: len(g)]


def preview_render(g, cam, bg=(0.0, 0.0, 0.0)):
    """GpuRenderer::render (gpu/compute.rs:210-218): (h, w, 3) linear RGB."""
    g = np.ascontiguousarray(g, np.float32).reshape(-1, 59)
    out = np.zeros((cam.height, cam.width, 3), np.float32)
    bgv = np.ascontiguousarray(bg, np.float32)
    c = camera(cam)
    if lib().orc_preview_render(_p(g), len(g), ctypes.byref(c), _p(bgv), _p(out)):
        raise ValueError("singular camera")
    return out


def resize_rgba8(img, width, height):
    """load_target's resize_exact(width, height, Triangle) (io/images.rs:57-61,
    image 0.25.10) of an (h, w, 4) RGBA8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros((height, width, 4), np.uint8)
    if lib().orc_resize_rgba8(_p(img), img.shape[1], img.shape[0], _p(out), width, height):
        raise ValueError("zero image size")
    return out
