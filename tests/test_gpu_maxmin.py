"""GPU tracer_maxmin (check_nans.F90:164-203) through the C-ABI against the oracle: each
tracer's maximum and minimum over the prognostic levels, exactly equal on uploaded
fields (with values planted in level 1 that must not show), and after a tracer step
within the step's own parity tolerance (F32 1e-5, F64 1e-12 relative)."""

import numpy as np
import pytest

from olam_amd import abi
from olam_amd.gpu import Context
from helpers import tracer_case
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tracer_maxmin_parity(grid_factory, prec):
    grid = grid_factory(12, 16)
    w, st, tr = tracer_case(grid, ntr=4)
    rng = np.random.default_rng(8)
    tr[3] = (rng.standard_normal((grid.mwa, grid.mza)) - 0.5).astype(np.float32)
    tr[3][:, 0] = 50.0                       # level 1 is not prognostic
    prm = abi.make_params(prec, num_scalar=4, dtsm=600.0)
    print(f"tracer_maxmin {prec}", flush=True)
    res = {}
    for name, be in (("gpu", Context(grid, prm, device=0)), ("oracle", Oracle(grid, prm))):
        be.upload_state(st)
        for n, q in enumerate(tr, start=1):
            be.upload_scalar(n, q)
        before = be.tracer_maxmin()
        be.scalar_transport_rk(w["rho_old"])
        res[name] = (before, be.tracer_maxmin())
        be.close()
    assert np.array_equal(res["gpu"][0], res["oracle"][0])
    tol = 1e-5 if prec == "f32" else 1e-12
    a, b = res["gpu"][1].astype(np.float64), res["oracle"][1].astype(np.float64)
    assert np.all(np.abs(a - b) <= tol * np.abs(b)), (a, b)
    assert res["gpu"][0][3, 0] < 50.0
