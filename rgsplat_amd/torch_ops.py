"""torch.autograd.Function over the HIP renderer.

Forward is render_color (cpu/backward.rs:152-166) on device tensors; backward
is the raster backward (backward.rs:391-512) plus geometry backprop
(:518-638) driven by the incoming dL/dRGB, so any torch loss can sit on top of
the renderer. Gradients come back in Gaussian3D layout, (n, 59): the
rotation gradient has no radial component (quaternion.rs:45-70), exactly as
in the reference.

The library's forward state is single-slot, so backward re-runs the forward
of its own inputs first; that keeps interleaved forwards (as in gradcheck)
correct.

torch ships its own HIP runtime (libamdhip64.so.7). Import torch before the
first use of rgsplat_amd.runtime in a process, so that the library binds to
the runtime torch already loaded instead of a second copy.
"""
from __future__ import annotations

import numpy as np
import torch

from . import runtime

# Gaussian3D offset -> GaussianGrad / flat index
_G3D_FROM_FLAT = [0, 1, 2, 10, 3, 4, 5, 6, 7, 8, 9] + list(range(11, 59))


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gaussians, cam, bg, rctx):
        if gaussians.dtype != torch.float32 or not gaussians.is_cuda:
            raise TypeError("gaussians must be a float32 device tensor of shape (n, 59)")
        g = gaussians.contiguous()
        rgb = torch.empty((cam.height, cam.width, 3), dtype=torch.float32, device=g.device)
        _forward(rctx, g, cam, bg, rgb)
        ctx.save_for_backward(g)
        ctx.cam, ctx.bg, ctx.rctx = cam, bg, rctx
        return rgb

    @staticmethod
    def backward(ctx, grad_rgb):
        (g,) = ctx.saved_tensors
        rctx = ctx.rctx
        rgb = torch.empty((ctx.cam.height, ctx.cam.width, 3), dtype=torch.float32, device=g.device)
        _forward(rctx, g, ctx.cam, ctx.bg, rgb)
        dl = grad_rgb.contiguous().to(torch.float32)
        grads = torch.empty((g.shape[0], 60), dtype=torch.float32, device=g.device)
        stream = torch.cuda.current_stream(g.device).cuda_stream
        runtime._check(runtime.lib().rgs_set_stream(rctx.h, stream))
        runtime._check(runtime.lib().rgs_backward_dev(rctx.h, dl.data_ptr(), grads.data_ptr()))
        return grads[:, _G3D_FROM_FLAT], None, None, None


def _forward(rctx, g, cam, bg, rgb):
    stream = torch.cuda.current_stream(g.device).cuda_stream
    runtime._check(runtime.lib().rgs_set_stream(rctx.h, stream))
    c = runtime.camera(cam)
    bgv = np.ascontiguousarray(bg, np.float32)
    import ctypes
    runtime._check(runtime.lib().rgs_forward_dev(rctx.h, g.data_ptr(), g.shape[0], ctypes.byref(c),
                                                 bgv.ctypes.data_as(ctypes.c_void_p), rgb.data_ptr()))


def render(gaussians: torch.Tensor, cam, bg=(0.0, 0.0, 0.0), rctx: runtime.Context | None = None) -> torch.Tensor:
    """Differentiable render of (n, 59) Gaussian3D records -> (h, w, 3)."""
    if rctx is None:
        rctx = _default_context(gaussians.device.index or 0)
    return _Render.apply(gaussians, cam, tuple(float(x) for x in bg), rctx)


_contexts: dict = {}


def _default_context(device: int) -> runtime.Context:
    if device not in _contexts:
        _contexts[device] = runtime.Context(device)
    return _contexts[device]
