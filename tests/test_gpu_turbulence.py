"""GPU parity of the eddy and horizontal diffusion in prog_wrtv_rk's setup loop
(solve_eddy_diff_heat / _vxe, pbl_driver.f90:519-742; comp_horiz_k, pbl_driver.f90:
136-238; prog_wrtv_rk.f90:213-251) against the oracle, through the C-ABI, with and
without the akmin background diffusivity; and zero turbulence fields leave the step
bit-identical.  Tolerances as in test_gpu_parity.py except the FP64 vertical fluxes
(below); also at L=70 (tall columns: two waves per column, each right-hand side of the
momentum diffusion solved in turn).
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from olam_amd.gpu import Context
from helpers import DYN_FIELDS, dyn_fields, rel_linf
from oracle import Oracle
from test_turbulence import turbulence

pytestmark = pytest.mark.gpu

DT = 60.0


def _case(nlev=20):
    grid = abi.FortranGrid(M.build_mesh(12), M.vertical_grid(nlev))
    st = cases.jet_state(grid)
    rng = np.random.default_rng(21)
    st.arrays["thil"][1:] += rng.standard_normal(st.arrays["thil"][1:].shape).astype(np.float32)
    st.arrays["thilt"][1:] = 1e-4 * rng.standard_normal(st.arrays["thilt"][1:].shape).astype(np.float32)
    return grid, st


@pytest.mark.parametrize("prec,akmin,nlev", [("f64", 0.0, 20), ("f32", 0.0, 20), ("f64", 0.3, 20), ("f32", 0.3, 20),
                                             ("f64", 0.0, 70), ("f32", 0.3, 70)])
def test_turbulence_step_parity(prec, akmin, nlev):
    grid, st = _case(nlev)
    turb = turbulence(grid, st, seed=5, drag=True)
    prm = abi.make_params(prec, num_scalar=1, dtsm=DT, akmin=akmin)
    print(f"turbulence {prec} akmin {akmin} L={nlev}", flush=True)
    out = {}
    for name, be in (("gpu", Context(grid, prm, device=0)), ("oracle", Oracle(grid, prm))):
        be.upload_state(st)
        be.init_wrtv_rk()
        be.set_turbulence(*turb)
        be.tend0()
        for _ in range(3):
            be.prog_wrtv_rk()
        out[name] = dyn_fields(be, grid)
        be.close()
    tol_big = 1e-11 if prec == "f64" else 1e-5
    # wmc, wc, wmasc carry the W solve's amplified rounding (FP64 PCR on the GPU, Thomas
    # in the oracle) and come back through the FP32 state download: one flipped last
    # bit at a value near the maximum is 2**-23 relative
    tol_small = 1.2e-7 if prec == "f64" else 1e-4
    for f in DYN_FIELDS:
        e = rel_linf(out["gpu"][f], out["oracle"][f])
        print(f"  {f:6s} rel Linf {e:.3e}")
        tol = tol_small if f in ("wmc", "wc", "wmasc") else tol_big
        if prec == "f64" and f == "wc":
            tol = 1e-7
        if prec == "f64" and nlev > 63 and f == "vze":
            tol = 1e-9     # tall columns: wc's PCR differences reach vze (test_gpu_tall_columns.py)
        assert e < tol, (f, e)


def test_zero_turbulence_is_a_no_op():
    grid, st = _case()
    z = np.zeros((grid.mwa, grid.mza), np.float32)
    zi = np.zeros(grid.mwa, np.int32)
    prm = abi.make_params("f32", num_scalar=1, dtsm=DT)
    res = []
    for turb in (None, (z, z, zi, zi, np.zeros(grid.mwa, np.float32))):
        g = Context(grid, prm, device=0)
        g.upload_state(st)
        g.init_wrtv_rk()
        if turb is not None:
            g.set_turbulence(*turb)
        g.tend0()
        g.prog_wrtv_rk()
        res.append(dyn_fields(g, grid))
        g.close()
    for f in DYN_FIELDS:
        assert np.array_equal(res[0][f], res[1][f]), f


def test_turbulence_input_checks():
    grid, st = _case()
    prm = abi.make_params("f32", num_scalar=1, dtsm=DT)
    g = Context(grid, prm, device=0)
    vkh, vkm, khtop, kmtop, akm = turbulence(grid, st)
    bad = khtop.copy()
    bad[3] = grid.mza
    with pytest.raises(RuntimeError, match="khtop"):
        g.set_turbulence(vkh, vkm, bad, kmtop, akm)
    g.close()
