"""GPU parity of the monotone θ/wind fluxes of the final acoustic RK stage
(fluxes_wrtv_monot, prog_wrtv_rk.f90:1133-1554; ithil_monot / iwind_monot > 0) against
the oracle, through the C-ABI, at horizontal orders 2 and 3, plus exact ρθ_il
conservation on the GPU (each cell rebuilds its face values itself, so a face whose two
cells disagree would show here).  Tolerances as for the dynamics step in
test_gpu_parity.py; also at L=70 (tall columns, two waves per column).
"""

import numpy as np
import pytest
import torch

from olam_amd import abi, cases
from olam_amd.gpu import Context
from helpers import DYN_FIELDS, dyn_fields, rel_linf
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _step_state(grid, amp=20.0):
    st = cases.jet_state(grid)
    lat, lon = cases._latlon_w(grid.mesh)
    zt = grid.vg["zt"]
    sec = ((lon % (2 * np.pi)) < 0.6 * np.pi) & (np.abs(lat) < 0.8)
    th = st.arrays["thil"]
    for k in range(2, grid.mza + 1):
        if 3000.0 < zt[k] < 8000.0:
            th[1:, k - 1] += np.where(sec, amp, 0.0).astype(np.float32)
    return st


@pytest.mark.parametrize("prec,order,nlev", [("f64", 2, 20), ("f32", 2, 20), ("f64", 3, 20), ("f32", 3, 20),
                                             ("f64", 2, 70), ("f32", 2, 70)])
def test_monotone_fluxes_parity(grid_factory, prec, order, nlev):
    grid = grid_factory(12, nlev)
    st = _step_state(grid)
    prm = abi.make_params(prec, num_scalar=1, dtsm=60.0, order=order, ithil_monot=1, iwind_monot=1)
    print(f"monotone fluxes {prec} order {order} L={nlev}", flush=True)
    g, o = Context(grid, prm, device=0), Oracle(grid, prm)
    out = {}
    for name, be in (("gpu", g), ("oracle", o)):
        be.upload_state(st)
        be.tend0()
        for _ in range(2):
            be.prog_wrtv_rk()
        out[name] = dyn_fields(be, grid)
    tol_big = 1e-11 if prec == "f64" else 1e-5
    tol_small = 5e-9 if prec == "f64" else 1e-4
    for f in DYN_FIELDS:
        e = rel_linf(out["gpu"][f], out["oracle"][f])
        print(f"  {f:6s} rel Linf {e:.3e}")
        tol = tol_small if f in ("wmc", "wc", "wmasc") else tol_big
        if prec == "f64" and f == "wc":
            tol = 1e-7
        if prec == "f64" and nlev > 63 and f in ("wmc", "wc", "wmasc", "vze"):
            # tall columns (two waves per column): the 128-row PCR against the oracle's
            # Thomas sweep, as in test_gpu_tall_columns.py
            tol = 1e-7
        assert e < tol, (f, e)
    g.close()


def test_monotone_fluxes_conserve_rho_thil_on_gpu(grid_factory):
    grid = grid_factory(12, 20)
    st = _step_state(grid)
    prm = abi.make_params("f64", num_scalar=1, dtsm=60.0, ithil_monot=1, iwind_monot=1)
    g = Context(grid, prm, device=0)
    g.upload_state(st)
    g.tend0()
    vol = torch.tensor(grid.arrays["volt"][1:, 1:], dtype=torch.float64, device="cuda")

    def rt():
        rho = torch.empty((grid.mwa - 1, grid.mza), dtype=torch.float64, device="cuda")
        th = torch.empty_like(rho)
        torch.cuda.synchronize()
        g.copy_field("rho", 0, rho.data_ptr(), 0)
        g.copy_field("thil", 0, th.data_ptr(), 0)
        torch.cuda.synchronize()
        return float((rho[:, 1:] * th[:, 1:] * vol).sum())

    a = rt()
    for _ in range(3):
        g.prog_wrtv_rk()
    b = rt()
    g.close()
    print(f"GPU rho*thil change over 3 monotone steps {abs(b / a - 1):.3e}")
    assert abs(b / a - 1) < 1e-13
