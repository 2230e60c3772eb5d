"""Context creation when the device cannot be used. The reference falls back
to the CPU when its GPU context fails to start (pipeline/mod.rs:72-86); this
library has no CPU path, so rgs_create fails with a status code and a message
instead, and the caller decides."""
import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rgsplat_amd", "librgsplat_hip.so")

PROBE = r"""
import ctypes, sys
L = ctypes.CDLL(sys.argv[1])
L.rgs_last_error.restype = ctypes.c_char_p
h = ctypes.c_void_p()
for dev in (0, -1, 7):
    rc = L.rgs_create(dev, ctypes.byref(h))
    print(dev, rc, (L.rgs_last_error() or b"").decode(), bool(h.value))
"""


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present (tests/test_gpu_parity.py covers it)")
def test_create_without_a_device_fails_cleanly():
    r = subprocess.run([sys.executable, "-c", PROBE, LIB], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for line in r.stdout.strip().splitlines():
        dev, rc, rest = line.split(" ", 2)
        msg, made = rest.rsplit(" ", 1)
        assert rc == "2" and "no HIP device" in msg and made == "False", line  # RGS_E_HIP, no context
