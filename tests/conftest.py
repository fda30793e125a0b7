import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("mb3d-rs_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def hip():
    """The HIP backend on device 0 (GPU tests only)."""
    import mb3d_hip as H

    lib = H.load()
    H.check(lib.mb3d_hip_init(0), "mb3d_hip_init")
    return H
