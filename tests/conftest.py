import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU runs")


def _ensure_host_libs():
    """Host-only libraries are cheap to build; build them when missing so the
    CPU suite runs from a clean tree.  The HIP library is never built here."""
    from traccc_amd import build
    if not os.path.exists(os.path.join(build.LIB_DIR, "libtraccc_amd_sim.so")):
        build.build_sim()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_ckf.so")):
        build.build_oracle()


_ensure_host_libs()


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return True
