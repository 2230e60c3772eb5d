"""The device expf restatement (rgsplat_amd/csrc/rgs_math.h), compiled for
the host, equals glibc expf -- the libm call behind the reference's f32::exp
-- on every one of the 2^32 inputs."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_expf_restatement_exhaustive():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "expf_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp",
                        "-I", os.path.join(ROOT, "rgsplat_amd", "csrc"),
                        os.path.join(ROOT, "tests", "native", "expf_exhaustive.cpp"), "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "0", f"{out} mismatches against glibc expf"
