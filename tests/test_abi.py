"""The C-ABI library loads and exports every function the public header
declares (no GPU needed: nothing here calls into the device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rgsplat_hip.h")
LIB = os.path.join(ROOT, "rgsplat_amd", "librgsplat_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["rgs_render_and_backward", "rgs_render_color", "rgs_rasterize_rgba", "rgs_loss_grad",
                 "rgs_adam_step", "rgs_densify_and_prune", "rgs_reset_opacities", "rgs_train_step"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_the_header():
    from rgsplat_amd import runtime
    assert set(declared_functions()) == set(runtime.SIGNATURES)


def test_record_layouts():
    from rgsplat_amd import runtime
    assert ctypes.sizeof(runtime.Camera) == 100
    assert ctypes.sizeof(runtime.TrainCfg) == 72
