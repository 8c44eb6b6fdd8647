import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libifea_amd.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def c1_mesh():
    from impressionfea_amd.config import upset_small
    from impressionfea_amd.mesh import kuhn_cylinder
    wl = upset_small()
    x, t, ox, oy = kuhn_cylinder(wl.diameter, wl.height, wl.nxy, wl.nz, seed=wl.seed)
    return wl, x, t, ox, oy
