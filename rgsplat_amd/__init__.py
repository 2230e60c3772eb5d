"""rgsplat_amd -- MI355X (gfx950) training step for differentiable Gaussian
splatting, a drop-in for the rgsplat crate's Rayon CPU path.

The compute lives in hand-written HIP kernels in librgsplat_hip.so behind the
C ABI declared in include/rgsplat_hip.h; `runtime` mirrors the crate's render,
backward and optimizer API over that ABI and `torch_ops` exposes the renderer
as a torch.autograd.Function.
"""
from . import synthetic  # noqa: F401

__all__ = ["synthetic", "runtime"]


def __getattr__(name):
    # runtime/torch_ops load the HIP library; import them lazily so that the
    # synthetic generator is usable without it
    if name in ("runtime", "torch_ops"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
