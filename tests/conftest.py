import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "owt-krylov_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import owtk

    if owtk.device_count() == 0:
        pytest.fail("gpu test on a machine without a HIP device")
    c = owtk.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def c0():
    """64 x 32 triangulated grid, block 3, seed 0, seeded permutation + RCM."""
    import owtk

    return owtk.generate_split(64, 32, 3, 0, ordering=owtk.ORDER_RCM)
