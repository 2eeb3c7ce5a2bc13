import os
import sys

import pytest
# torch before libolamgpu in every test process: torch's HIP runtime then serves both
# (loaded after the library, torch would start a second runtime that finds no GPU)
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


_cache = {}


@pytest.fixture(scope="session")
def grid_factory():
    """Cached FortranGrid builder: grid_factory(nxp, nlev, order='natural')."""
    from olam_amd import abi, mesh

    def make(nxp, nlev, order="natural", seed=0):
        key = (nxp, nlev, order, seed)
        if key not in _cache:
            m = mesh.build_mesh(nxp, order=order, seed=seed)
            vg = mesh.vertical_grid(nlev)
            _cache[key] = abi.FortranGrid(m, vg)
        return _cache[key]

    return make
