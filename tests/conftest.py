import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: longer CPU test")
    # torch bundles its own HIP runtime; a process that uses torch tensors
    # with librgsplat_hip must load torch's runtime first so both share it
    # (every run that may select GPU tests: -m gpu, no -m, named files)
    markexpr = config.getoption("-m") or ""
    if "not gpu" not in markexpr:
        try:
            import torch  # noqa: F401
        except Exception:
            pass


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
