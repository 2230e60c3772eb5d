"""ctypes binding of librgsplat_hip.so and a Python mirror of the crate API.

The mirrored functions keep the reference's names, argument meaning and
layouts (cpu/backward.rs, cpu/rasterizer.rs, pipeline/optimizer.rs):
Gaussians are (n, 59) float32 Gaussian3D records, gradients (n, 60)
GaussianGrad records, images row-major (h, w, 3) RGB / (h, w, 4) RGBA float32.
Failures raise RgsError with the library's message instead of panicking.

There is no CPU fallback: importing this module without the built library, or
calling it on a machine without a HIP device, raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import synthetic

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RGS_LIB", os.path.join(HERE, "librgsplat_hip.so"))

_lib = None


class RgsError(RuntimeError):
    pass


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


class AdamRec(ctypes.Structure):
    _fields_ = [("m1", ctypes.c_void_p), ("m2", ctypes.c_void_p), ("grad_accum", ctypes.c_void_p),
                ("grad_denom", ctypes.c_void_p), ("rows", ctypes.c_size_t), ("capacity", ctypes.c_size_t),
                ("t", ctypes.c_uint64)]


# exported symbol -> argtypes (restype int unless noted)
_vp, _sz, _i, _f = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_float
_CAM = ctypes.POINTER(Camera)
_CFG = ctypes.POINTER(TrainCfg)
_ADAM = ctypes.POINTER(AdamRec)
SIGNATURES = {
    "rgs_create": [_i, ctypes.POINTER(_vp)],
    "rgs_destroy": [_vp],
    "rgs_last_error": [],
    "rgs_set_stream": [_vp, _vp],
    "rgs_use_own_stream": [_vp],
    "rgs_render_and_backward": [_vp, _vp, _sz, _CAM, _vp, _vp, _f, ctypes.POINTER(_f), _vp],
    "rgs_render_color": [_vp, _vp, _sz, _CAM, _vp, _vp],
    "rgs_rasterize_rgba": [_vp, _vp, _sz, _CAM, _vp, _vp],
    "rgs_loss_grad": [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _f, ctypes.POINTER(_f), _vp],
    "rgs_adam_step": [_vp, _vp, _vp, _sz, _ADAM, _CFG],
    "rgs_densify_and_prune": [_vp, _vp, ctypes.POINTER(_sz), _sz, _ADAM, _CFG, _f],
    "rgs_reset_opacities": [_vp, _vp, _sz],
    "rgs_forward_dev": [_vp, _vp, _sz, _CAM, _vp, _vp],
    "rgs_backward_dev": [_vp, _vp, _vp],
    "rgs_scene_upload": [_vp, _vp, _sz, _sz],
    "rgs_scene_size": [_vp, ctypes.POINTER(_sz)],
    "rgs_scene_download": [_vp, _vp, _sz, ctypes.POINTER(_sz)],
    "rgs_set_views_rgba8": [_vp, _vp, _vp, _i],
    "rgs_rgba8_to_linear": [_vp, _vp, _sz, _vp],
    "rgs_to_u8_srgb": [_vp, _vp, _sz, _vp],
    "rgs_save_ply": [_vp, _vp, _sz, ctypes.c_char_p],
    "rgs_scene_save_ply": [_vp, ctypes.c_char_p],
    "rgs_load_ply": [_vp, ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_sz)],
    "rgs_scene_load_ply": [_vp, ctypes.c_char_p, _sz],
    "rgs_save_splat": [_vp, _vp, _sz, ctypes.c_char_p],
    "rgs_scene_save_splat": [_vp, ctypes.c_char_p],
    "rgs_resize_rgba8": [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32, ctypes.c_uint32],
    "rgs_set_views_rgba8_sized": [_vp, _CAM, ctypes.POINTER(ctypes.c_void_p), _vp, _i],
    "rgs_preview_project": [_vp, _vp, _sz, _CAM, _vp],
    "rgs_preview_render": [_vp, _vp, _sz, _CAM, _vp, _vp],
    "rgs_scene_preview_render": [_vp, _CAM, _vp, _vp],
    "rgs_scene_preview_render_device": [_vp, _CAM, _vp, _vp],
    "rgs_gaussians_from_sfm": [_vp, _vp, _vp, _sz, _vp],
    "rgs_scene_from_sfm": [_vp, _vp, _vp, _sz, _sz],
    "rgs_adam_download": [_vp, _ADAM],
    "rgs_set_views": [_vp, _CAM, ctypes.POINTER(ctypes.c_void_p), _i],
    "rgs_train_step": [_vp, ctypes.POINTER(_i), _i, _CFG, _vp, _f, _i, _vp],
    "rgs_scene_densify": [_vp, _CFG, _f],
    "rgs_scene_reset_opacities": [_vp],
    "rgs_step_stats": [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                       ctypes.POINTER(ctypes.c_uint64)],
    "rgs_set_depth_cut": [_vp, _i, _f],
    "rgs_set_train_binning": [_vp, _i, _i],
    "rgs_step_image": [_vp, _vp, _sz],
    "rgs_cut_stats": [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_float),
                      ctypes.POINTER(ctypes.c_uint64)],
    "rgs_set_raster_params": [_vp, _i, _i],
    "rgs_get_raster_params": [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "rgs_set_raster_batch": [_vp, _i],
    "rgs_get_raster_batch": [_vp, ctypes.POINTER(ctypes.c_int)],
    "rgs_set_raster_queue": [_vp, _i],
    "rgs_get_raster_queue": [_vp, ctypes.POINTER(ctypes.c_int)],
    "rgs_autotune_raster": [_vp, _vp, _i, _vp, ctypes.c_float, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)],
    "rgs_eval_expf": [_vp, _vp, _vp, _sz],
    "rgs_set_profiling": [_vp, _i],
    "rgs_set_profiling_stages": [_vp, ctypes.c_uint32],
    "rgs_stage_times": [_vp, _vp, _i, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "rgs_debug_tiles": [_vp, _vp, _sz, _vp, _sz, ctypes.POINTER(ctypes.c_uint64)],
    "rgs_set_grad_capture": [_vp, _i],
    "rgs_raster_work_counts": [_vp, _vp],
    "rgs_hbm_copy_probe": [_vp, ctypes.c_size_t, _i, ctypes.POINTER(ctypes.c_float)],
    "rgs_set_split_scale_divisor": [_vp, _f],
    "rgs_loss_grad_dev_noalloc": [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _f, _vp, _vp],
    "rgs_step_grads": [_vp, _vp, _sz],
}


def lib():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m rgsplat_amd.build` "
                              "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.rgs_destroy.restype = None
        L.rgs_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int):
    if rc != 0:
        msg = lib().rgs_last_error()
        raise RgsError(f"rgsplat_hip error {rc}: {msg.decode() if msg else ''}")


def camera(cam) -> Camera:
    c = Camera()
    c.width, c.height = int(cam.width), int(cam.height)
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    for i, v in enumerate(np.asarray(cam.c2w, dtype=np.float32).reshape(16)):
        c.c2w[i] = float(v)
    return c


def train_cfg(d: dict | None = None) -> TrainCfg:
    d = d or synthetic.train_config()
    c = TrainCfg()
    for k, _ in TrainCfg._fields_:
        setattr(c, k, d[k])
    return c


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class AdamState:
    """AdamState (optimizer.rs:74-106): m1/m2 rows x 59, grad statistics, t.
    Buffers are allocated with room for `capacity` rows."""

    def __init__(self, n: int, capacity: int | None = None):
        cap = max(capacity or n, n, 1)
        self.m1 = np.zeros((cap, 59), np.float32)
        self.m2 = np.zeros((cap, 59), np.float32)
        self.grad_accum = np.zeros(cap, np.float32)
        self.grad_denom = np.zeros(cap, np.uint32)
        self.rows = n
        self.t = 0

    def rec(self) -> AdamRec:
        r = AdamRec()
        r.m1, r.m2 = _p(self.m1), _p(self.m2)
        r.grad_accum, r.grad_denom = _p(self.grad_accum), _p(self.grad_denom)
        r.rows, r.capacity, r.t = self.rows, len(self.m1), self.t
        return r

    def pull(self, r: AdamRec):
        self.rows, self.t = r.rows, r.t


class Context:
    """One HIP device context (stream + device buffers)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().rgs_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().rgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------- reference API mirror
    def render_and_backward(self, gaussians, cam, target, bg=(0.0, 0.0, 0.0), lambda_dssim=0.2):
        """backward.rs:107-149 -> (loss, grads (n, 60))."""
        g = _f32(gaussians)
        n = len(g)
        tg = _f32(target)
        if tg.size != cam.width * cam.height * 4:
            raise RgsError("target must be h*w*4")
        grads = np.zeros((n, 60), np.float32)
        loss = ctypes.c_float()
        c = camera(cam)
        _check(lib().rgs_render_and_backward(self.h, _p(g), n, ctypes.byref(c), _p(tg), _p(_f32(bg)),
                                             float(lambda_dssim), ctypes.byref(loss), _p(grads)))
        return float(loss.value), grads

    def render_color(self, gaussians, cam, bg=(0.0, 0.0, 0.0)):
        """backward.rs:152-166 -> (h, w, 3)."""
        g = _f32(gaussians)
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        c = camera(cam)
        _check(lib().rgs_render_color(self.h, _p(g), len(g), ctypes.byref(c), _p(_f32(bg)), _p(out)))
        return out

    def rasterize(self, gaussians, cam, bg=(0.0, 0.0, 0.0)):
        """rasterizer.rs:72-192 -> Framebuffer data (h, w, 4)."""
        g = _f32(gaussians)
        out = np.zeros((cam.height, cam.width, 4), np.float32)
        c = camera(cam)
        _check(lib().rgs_rasterize_rgba(self.h, _p(g), len(g), ctypes.byref(c), _p(_f32(bg)), _p(out)))
        return out

    def preview_project(self, gaussians, cam):
        """GpuRenderer::project (gpu/compute.rs:138-207) -> (n, 10) SplatInfo2D."""
        g = _f32(gaussians).reshape(-1, 59)
        out = np.zeros((max(len(g), 1), 10), np.float32)
        c = camera(cam)
        _check(lib().rgs_preview_project(self.h, _p(g), len(g), ctypes.byref(c), _p(out)))
        return out[: len(g)]

    def preview_render(self, gaussians, cam, bg=(0.0, 0.0, 0.0)):
        """GpuRenderer::render (gpu/compute.rs:210-218) -> (h, w, 3) linear RGB."""
        g = _f32(gaussians).reshape(-1, 59)
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        c = camera(cam)
        _check(lib().rgs_preview_render(self.h, _p(g), len(g), ctypes.byref(c), _p(_f32(bg)), _p(out)))
        return out

    def scene_preview_render(self, cam, bg=(0.0, 0.0, 0.0)):
        """The preview render of the device scene."""
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        c = camera(cam)
        _check(lib().rgs_scene_preview_render(self.h, ctypes.byref(c), _p(_f32(bg)), _p(out)))
        return out

    def scene_preview_render_device(self, cam, out, bg=(0.0, 0.0, 0.0)):
        """The preview render of the device scene into `out`, a contiguous
        float32 torch tensor of shape (h, w, 3) on this context's device."""
        if tuple(out.shape) != (cam.height, cam.width, 3) or not out.is_contiguous():
            raise ValueError("out must be a contiguous (h, w, 3) tensor")
        if str(out.dtype) != "torch.float32" or out.device.type != "cuda":
            raise ValueError("out must be a float32 tensor on the GPU")
        c = camera(cam)
        # The render runs on this context's stream, which is not ordered
        # after torch's: wait for the work torch has queued (which may still
        # write or read `out`). The call returns with the image complete.
        import torch
        torch.cuda.current_stream(out.device).synchronize()
        _check(lib().rgs_scene_preview_render_device(self.h, ctypes.byref(c), _p(_f32(bg)),
                                                     ctypes.c_void_p(out.data_ptr())))
        return out

    def photometric_loss_grad(self, rendered, target_rgba, w, h, lambda_dssim=0.2):
        """backward.rs:646-693 -> (loss, dL/drendered (w*h*3,))."""
        r = _f32(rendered)
        t = _f32(target_rgba)
        if r.size != w * h * 3 or t.size != w * h * 4:
            raise RgsError("rendered must be w*h*3 and target w*h*4")
        dl = np.zeros(w * h * 3, np.float32)
        loss = ctypes.c_float()
        _check(lib().rgs_loss_grad(self.h, _p(r), _p(t), w, h, float(lambda_dssim), ctypes.byref(loss), _p(dl)))
        return float(loss.value), dl

    def set_stream(self, hip_stream: int | None):
        """Run later work on an external HIP stream handle (e.g. a torch
        stream's cuda_stream); None returns to the context's own stream."""
        if hip_stream is None:
            _check(lib().rgs_use_own_stream(self.h))
        else:
            _check(lib().rgs_set_stream(self.h, ctypes.c_void_p(int(hip_stream))))

    def loss_grad_dev_noalloc(self, rgb_dev_ptr: int, target_dev_ptr: int, w: int, h: int, lambda_dssim: float,
                              loss_dev_ptr: int, dl_dev_ptr: int):
        """photometric_loss_grad on device pointers, allocating nothing
        (rgs_loss_grad_dev_noalloc); enqueued on the context's stream."""
        _check(lib().rgs_loss_grad_dev_noalloc(self.h, ctypes.c_void_p(rgb_dev_ptr), ctypes.c_void_p(target_dev_ptr),
                                               int(w), int(h), float(lambda_dssim), ctypes.c_void_p(loss_dev_ptr),
                                               ctypes.c_void_p(dl_dev_ptr)))

    def adam_step(self, gaussians: np.ndarray, grads, state: AdamState, cfg: dict | None = None):
        """optimizer.rs:116-163; updates `gaussians` (float32, C-contiguous) in place."""
        assert gaussians.dtype == np.float32 and gaussians.flags.c_contiguous
        gr = _f32(grads)
        rec = state.rec()
        c = train_cfg(cfg)
        _check(lib().rgs_adam_step(self.h, _p(gaussians), _p(gr), len(gaussians), ctypes.byref(rec),
                                   ctypes.byref(c)))
        state.pull(rec)

    def densify_and_prune(self, gbuf: np.ndarray, n: int, state: AdamState, cfg: dict | None, extent: float):
        """optimizer.rs:206-272 on a (capacity, 59) buffer whose first n rows
        are live; returns the new n."""
        assert gbuf.dtype == np.float32 and gbuf.flags.c_contiguous
        nn = ctypes.c_size_t(n)
        rec = state.rec()
        c = train_cfg(cfg)
        _check(lib().rgs_densify_and_prune(self.h, _p(gbuf), ctypes.byref(nn), len(gbuf), ctypes.byref(rec),
                                           ctypes.byref(c), float(extent)))
        state.pull(rec)
        return nn.value

    def reset_opacities(self, gaussians: np.ndarray):
        """optimizer.rs:314-319 (in place)."""
        assert gaussians.dtype == np.float32 and gaussians.flags.c_contiguous
        _check(lib().rgs_reset_opacities(self.h, _p(gaussians), len(gaussians)))

    # --------------------------------------------- device-resident training
    def scene_upload(self, gaussians, capacity: int | None = None):
        g = _f32(gaussians)
        _check(lib().rgs_scene_upload(self.h, _p(g), len(g), int(capacity or len(g))))

    def scene_size(self) -> int:
        n = ctypes.c_size_t()
        _check(lib().rgs_scene_size(self.h, ctypes.byref(n)))
        return n.value

    def scene_download(self) -> np.ndarray:
        n = self.scene_size()
        out = np.zeros((max(n, 1), 59), np.float32)
        nn = ctypes.c_size_t()
        _check(lib().rgs_scene_download(self.h, _p(out), len(out), ctypes.byref(nn)))
        return out[: nn.value]

    # ------------------------------------------------------ colour conversions
    def rgba8_to_linear(self, img) -> np.ndarray:
        """load_target's conversion of decoded RGBA8 (h, w, 4) to linear f32."""
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros(img.shape, np.float32)
        _check(lib().rgs_rgba8_to_linear(self.h, _p(img), img.size // 4, _p(out)))
        return out

    def to_u8_srgb(self, rgba) -> np.ndarray:
        """Framebuffer::to_u8_srgb of an RGBA f32 (h, w, 4) image."""
        rgba = _f32(rgba)
        out = np.zeros(rgba.shape, np.uint8)
        _check(lib().rgs_to_u8_srgb(self.h, _p(rgba), rgba.size // 4, _p(out)))
        return out

    def set_views_rgba8(self, cams, images):
        """set_views from decoded RGBA8 images of the cameras' sizes."""
        cs = (Camera * len(cams))(*[camera(c) for c in cams])
        ims = [np.ascontiguousarray(im, np.uint8) for im in images]
        for c, im in zip(cams, ims):
            if im.shape != (c.height, c.width, 4):
                raise ValueError(f"image {im.shape} does not match camera {(c.height, c.width, 4)}")
        ptrs = (ctypes.c_void_p * len(ims))(*[im.ctypes.data for im in ims])
        _check(lib().rgs_set_views_rgba8(self.h, cs, ptrs, len(ims)))
        self._views_keepalive = ims

    def set_views_rgba8_sized(self, cams, images):
        """set_views from decoded RGBA8 images (h, w, 4) of any size, resized
        on the device to the cameras' sizes as load_target does."""
        cs = (Camera * len(cams))(*[camera(c) for c in cams])
        ims = [np.ascontiguousarray(im, np.uint8) for im in images]
        for im in ims:
            if im.ndim != 3 or im.shape[2] != 4:
                raise ValueError(f"image {im.shape} is not (h, w, 4)")
        sizes = np.array([[im.shape[1], im.shape[0]] for im in ims], np.uint32)
        ptrs = (ctypes.c_void_p * len(ims))(*[im.ctypes.data for im in ims])
        _check(lib().rgs_set_views_rgba8_sized(self.h, cs, ptrs, _p(sizes), len(ims)))
        self._views_keepalive = ims

    def resize_rgba8(self, img, width, height):
        """image 0.25's resize_exact(width, height, Triangle) of (h, w, 4) RGBA8."""
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros((height, width, 4), np.uint8)
        _check(lib().rgs_resize_rgba8(self.h, _p(img), img.shape[1], img.shape[0], _p(out), width, height))
        return out

    # ------------------------------------------------------------ model files
    def save_ply(self, gaussians, path):
        """save_ply (io/ply.rs:35-102) of (n, 59) Gaussian3D records."""
        g = _f32(gaussians).reshape(-1, 59)
        _check(lib().rgs_save_ply(self.h, _p(g), len(g), os.fsencode(path)))

    def scene_save_ply(self, path):
        """The device scene as a PLY (save_ply's file)."""
        _check(lib().rgs_scene_save_ply(self.h, os.fsencode(path)))

    def load_ply(self, path) -> np.ndarray:
        """load_ply (io/ply.rs:116-247): (n, 59) Gaussian3D records."""
        n = ctypes.c_size_t()
        rc = lib().rgs_load_ply(self.h, os.fsencode(path), None, 0, ctypes.byref(n))
        if rc not in (0, 5):  # 5: RGS_E_CAPACITY, the count is in n
            _check(rc)
        out = np.zeros((max(n.value, 1), 59), np.float32)
        _check(lib().rgs_load_ply(self.h, os.fsencode(path), _p(out), len(out), ctypes.byref(n)))
        return out[: n.value]

    def scene_load_ply(self, path, capacity: int = 0):
        """A PLY into the device scene (as scene_upload)."""
        _check(lib().rgs_scene_load_ply(self.h, os.fsencode(path), int(capacity)))

    def gaussians_from_sfm(self, xyz, rgb) -> np.ndarray:
        """gaussians_from_sfm (pipeline/sfm.rs:365-407) of n SfM points
        (xyz (n, 3) f64, rgb (n, 3) u8): (n, 59) Gaussian3D records."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        rgb = np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3)
        if len(xyz) != len(rgb):
            raise ValueError("xyz and rgb differ in length")
        out = np.zeros((max(len(xyz), 1), 59), np.float32)
        _check(lib().rgs_gaussians_from_sfm(self.h, _p(xyz), _p(rgb), len(xyz), _p(out)))
        return out[: len(xyz)]

    def scene_from_sfm(self, xyz, rgb, capacity: int = 0):
        """The Gaussians of gaussians_from_sfm straight into the device scene."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        rgb = np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3)
        if len(xyz) != len(rgb):
            raise ValueError("xyz and rgb differ in length")
        _check(lib().rgs_scene_from_sfm(self.h, _p(xyz), _p(rgb), len(xyz), int(capacity)))

    def save_splat(self, gaussians, path):
        """save_splat (io/ply.rs:258-303) of (n, 59) Gaussian3D records."""
        g = _f32(gaussians).reshape(-1, 59)
        _check(lib().rgs_save_splat(self.h, _p(g), len(g), os.fsencode(path)))

    def scene_save_splat(self, path):
        _check(lib().rgs_scene_save_splat(self.h, os.fsencode(path)))

    def adam_download(self, capacity: int | None = None) -> AdamState:
        n = self.scene_size()
        st = AdamState(0, capacity=max(capacity or 0, n, 1) * 2)
        rec = st.rec()
        _check(lib().rgs_adam_download(self.h, ctypes.byref(rec)))
        st.pull(rec)
        return st

    def set_views(self, cams, targets):
        arr = (Camera * len(cams))(*[camera(c) for c in cams])
        self._targets = [_f32(t) for t in targets]
        ptrs = (ctypes.c_void_p * len(cams))(*[t.ctypes.data for t in self._targets])
        _check(lib().rgs_set_views(self.h, arr, ptrs, len(cams)))

    def train_step(self, view_ids, cfg: dict | None = None, bg=(0.0, 0.0, 0.0), lambda_dssim=0.2,
                   use_graph=True):
        ids = (ctypes.c_int * len(view_ids))(*[int(v) for v in view_ids])
        loss = np.zeros(len(view_ids), np.float32)
        c = train_cfg(cfg)
        _check(lib().rgs_train_step(self.h, ids, len(view_ids), ctypes.byref(c), _p(_f32(bg)),
                                    float(lambda_dssim), int(bool(use_graph)), _p(loss)))
        return loss

    def scene_densify(self, cfg: dict | None, extent: float):
        c = train_cfg(cfg)
        _check(lib().rgs_scene_densify(self.h, ctypes.byref(c), float(extent)))

    def set_split_scale_divisor(self, divisor: float):
        """Densify's split: 0 (default) the reference's (s0 / sqrt 2 on axis
        0); d > 1 divides every scale of the children by d."""
        _check(lib().rgs_set_split_scale_divisor(self.h, float(divisor)))

    def scene_reset_opacities(self):
        _check(lib().rgs_scene_reset_opacities(self.h))

    def step_pairs(self) -> int:
        k = ctypes.c_uint64()
        _check(lib().rgs_step_stats(self.h, ctypes.byref(k), None, None))
        return k.value

    def step_work(self) -> tuple[int, int, int]:
        """(tile pairs, forward pair evaluations, backward pair evaluations)
        of the last render."""
        k, f, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().rgs_step_stats(self.h, ctypes.byref(k), ctypes.byref(f), ctypes.byref(b)))
        return k.value, f.value, b.value

    def set_train_binning(self, shift: int, fwd_tile_blocks: int = 1):
        """train_step bins per 2^shift x 2^shift tiles (0 or 1, default 1);
        forward workgroups per tile (1) or per bin (0)."""
        _check(lib().rgs_set_train_binning(self.h, int(shift), int(fwd_tile_blocks)))

    def step_image(self, n_pixels: int) -> np.ndarray:
        """The last train_step's rendered images, flat (n_pixels, 3)."""
        out = np.zeros((n_pixels, 3), np.float32)
        _check(lib().rgs_step_image(self.h, _p(out), out.size))
        return out

    def set_depth_cut(self, mode: int, depth: float = 0.0):
        """Depth-cut binning of train_step: 0 adaptive (default), 1 off
        (whole lists), 2 fixed camera-space depth."""
        _check(lib().rgs_set_depth_cut(self.h, int(mode), float(depth)))

    def cut_stats(self) -> dict:
        """Last step's front / back pair counts, workgroups that needed their
        back lists, the next step's cut depth (inf: none) and the steps re-run
        so far because a step without the back-list stage needed it."""
        a, b, t, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        d = ctypes.c_float()
        _check(lib().rgs_cut_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t), ctypes.byref(d),
                                   ctypes.byref(r)))
        return {"front_pairs": a.value, "back_pairs": b.value, "flagged_tiles": t.value, "next_cut": d.value,
                "reruns": r.value}

    def set_raster_params(self, super_tile: int, sparse_lanes: int):
        _check(lib().rgs_set_raster_params(self.h, int(super_tile), int(sparse_lanes)))

    def raster_params(self) -> tuple[int, int]:
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().rgs_get_raster_params(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_raster_batch(self, batch: int):
        """Splats the binned training raster stages per batch: 32, 64, 128."""
        _check(lib().rgs_set_raster_batch(self.h, int(batch)))

    def raster_batch(self) -> int:
        b = ctypes.c_int()
        _check(lib().rgs_get_raster_batch(self.h, ctypes.byref(b)))
        return b.value

    def set_raster_queue(self, mode: int):
        """Persistent tile queue of the training raster: bit 0 the forward,
        bit 1 the backward (0: one workgroup per bin)."""
        _check(lib().rgs_set_raster_queue(self.h, int(mode)))

    def raster_queue(self) -> int:
        q = ctypes.c_int()
        _check(lib().rgs_get_raster_queue(self.h, ctypes.byref(q)))
        return q.value

    def hbm_copy_probe(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """GB/s (read + written) of a plain streaming device copy of nbytes:
        the practical ceiling of the memory-bound kernels."""
        g = ctypes.c_float()
        _check(lib().rgs_hbm_copy_probe(self.h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    def autotune_raster(self, view_ids, bg=(0.0, 0.0, 0.0), lambda_dssim=0.2) -> dict:
        """Times the raster under candidate work decompositions on the
        registered views and keeps the fastest (rgs_autotune_raster)."""
        ids = np.ascontiguousarray(view_ids, np.int32)
        bgv = np.ascontiguousarray(bg, np.float32)
        a, b, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        _check(lib().rgs_autotune_raster(self.h, _p(ids), len(ids), _p(bgv), float(lambda_dssim), ctypes.byref(a),
                                         ctypes.byref(b), ctypes.byref(ms)))
        return {"super_tile": a.value, "sparse_lanes": b.value, "batch": self.raster_batch(),
                "queue": self.raster_queue(), "raster_ms": ms.value}

    # ---------------------------------------------------------- stage timing
    STAGES = ["project", "depth_sort", "pair_scan", "pair_gen", "tile_sort", "raster_fwd", "loss_ssim",
              "raster_bwd", "geometry_adam", "cut_back"]

    def set_profiling(self, enable: bool):
        _check(lib().rgs_set_profiling(self.h, int(bool(enable))))

    def set_profiling_stages(self, stages):
        """Stage events around the named stages only (names as in STAGES)."""
        mask = 0
        for name in stages:
            mask |= 1 << self.STAGES.index(name)
        _check(lib().rgs_set_profiling_stages(self.h, mask))

    def stage_times(self) -> tuple[dict, int, int]:
        """(mean ms per stage, profiled steps, of which graph replays)."""
        ms = np.zeros(len(self.STAGES), np.float32)
        n = ctypes.c_uint64()
        ng = ctypes.c_uint64()
        _check(lib().rgs_stage_times(self.h, _p(ms), len(ms), ctypes.byref(n), ctypes.byref(ng)))
        return dict(zip(self.STAGES, ms.tolist())), n.value, ng.value

    # ------------------------------------------------------------ diagnostics
    def eval_expf(self, x) -> np.ndarray:
        x = _f32(x).reshape(-1)
        out = np.zeros_like(x)
        _check(lib().rgs_eval_expf(self.h, _p(x), _p(out), len(x)))
        return out

    def raster_work_counts(self) -> dict:
        """The last train_step's raster work, counted by the kernels
        (rgs_raster_work_counts)."""
        out = np.zeros(6, np.uint64)
        _check(lib().rgs_raster_work_counts(self.h, _p(out)))
        keys = ["fwd_lane_evals", "fwd_wave_visits", "fwd_composites", "bwd_lane_evals", "bwd_wave_visits",
                "bwd_contributions"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_grad_capture(self, enable: bool):
        """train_step also keeps the gradients its Adam update consumed."""
        _check(lib().rgs_set_grad_capture(self.h, int(bool(enable))))

    def step_grads(self) -> np.ndarray:
        """The last train_step's gradients, (n, 60) GaussianGrad records
        summed over its views (needs set_grad_capture(True) before the step)."""
        n = self.scene_size()
        out = np.zeros((max(n, 1), 60), np.float32)
        _check(lib().rgs_step_grads(self.h, _p(out), n))
        return out[:n]

    def debug_tiles(self, n_tiles: int, k_max: int):
        off = np.zeros(n_tiles + 1, np.uint64)
        gis = np.zeros(max(k_max, 1), np.uint32)
        k = ctypes.c_uint64()
        _check(lib().rgs_debug_tiles(self.h, _p(off), len(off), _p(gis), len(gis), ctypes.byref(k)))
        return off, gis[: k.value]


def scene_extent(gaussians) -> float:
    """estimate_scene_extent (pipeline/mod.rs:299-313): bbox diagonal of the
    positions (f32 arithmetic)."""
    g = np.asarray(gaussians, np.float32)
    if len(g) == 0:
        return 1.0
    d = (g[:, 0:3].max(0) - g[:, 0:3].min(0)).astype(np.float32)
    return float(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
