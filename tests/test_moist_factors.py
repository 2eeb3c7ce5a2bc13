"""Oracle pins of comp_alpha_press (timestep.f90:500-559), the moist pressure and
pressure-gradient factors that prog_wrtv_rk reads, on the CPU.

* Dry limit: with no water and theta = thil the factors reduce to the dry constants
  alpha_press = pc1 * rdry**cpocv (mem_basic.f90:50), pwfac = dzim, pvfac = dniv *
  zfacit.
* Gas law: for a moist state built from (p, T, r_v) with rho = p / ((rdry + rvap r_v)
  T) and theta = T (p00 / p)**(rdry / cp), alpha_press * (rho * thil)**cpocv gives p
  back (the equation of state prog_wrt_begs uses for press).
* The dry-air fractions pwfac / dzim and pvfac / (dniv zfacit) approximate the mean of
  1 / (1 + r_w) over the two points to second order in r_w.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle


@pytest.fixture(scope="module")
def grid():
    return abi.FortranGrid(M.build_mesh(6), M.vertical_grid(16))


def _run(grid, st, rr_v, rr_w, theta, prec="f64"):
    o = Oracle(grid, abi.make_params(prec, dtsm=60.0))
    o.upload_state(st)
    o.comp_alpha_press(rr_v, rr_w, theta)
    out = abi.FortranState(grid)
    o.download_state(out)
    o.close()
    return out.arrays


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_dry_limit(grid, prec):
    st = cases.jet_state(grid)
    z = np.zeros((grid.mwa, grid.mza), np.float32)
    a = _run(grid, st, z, z, st.arrays["thil"], prec)
    mza = grid.mza
    # alpha_press comes back through the FP32 state download in both modes
    np.testing.assert_allclose(a["alpha_press"][1:, 1:], M.ALPHA_DRY, rtol=3e-7)
    dzim = grid.arrays["dzim"]
    assert np.array_equal(a["pwfac"][1:, 1:mza - 1], np.broadcast_to(dzim[1:mza - 1], (grid.mwa - 1, mza - 2)))
    dn = grid.arrays["dniv"][1:, None] * grid.arrays["zfacit"][None, 1:]
    assert np.array_equal(a["pvfac"][1:, 1:], dn.astype(np.float32))


def test_gas_law(grid):
    rng = np.random.default_rng(3)
    shape = (grid.mwa, grid.mza)
    p = rng.uniform(2.0e4, 1.0e5, shape)
    T = rng.uniform(200.0, 300.0, shape)
    rv = rng.uniform(0.0, 0.02, shape)
    rdry, rvap, cp, cv, p00 = 287.0, 461.0, 1004.0, 717.0, 1.0e5
    rho = p / ((rdry + rvap * rv) * T)
    theta = T * (p00 / p) ** (rdry / cp)
    st = abi.FortranState(grid)
    st.arrays["thil"][:] = theta.astype(np.float32)
    st.arrays["rho"][:] = rho
    a = _run(grid, st, rv.astype(np.float32), rv.astype(np.float32), st.arrays["thil"])
    # the oracle reads rr_v and theta in FP32: rebuild p from those inputs
    rv32 = rv.astype(np.float32).astype(np.float64)
    th32 = theta.astype(np.float32).astype(np.float64)
    T32 = th32 * (p / p00) ** (rdry / cp)
    p_eos = rho * (rdry + rvap * rv32) * T32
    p_mod = a["alpha_press"] * (rho * th32) ** (cp / cv)
    rel = np.abs(p_mod[1:, 1:] / p_eos[1:, 1:] - 1.0)
    # alpha_press (FP64 here) comes back through the FP32 state download
    assert rel.max() < 2e-7


def test_dry_fractions(grid):
    rng = np.random.default_rng(4)
    rw = rng.uniform(0.0, 0.03, (grid.mwa, grid.mza)).astype(np.float32)
    st = cases.jet_state(grid)
    a = _run(grid, st, rw, rw, st.arrays["thil"])
    mza = grid.mza
    r = rw.astype(np.float64)
    zt, zb = grid.vg["zwgt_top8"], grid.vg["zwgt_bot8"]
    exact_w = np.zeros_like(r)
    for k in range(2, mza):          # Fortran level k at column index k-1
        exact_w[:, k - 1] = zb[k] / (1.0 + r[:, k - 1]) + zt[k] / (1.0 + r[:, k])
    dzim = grid.arrays["dzim"].astype(np.float64)
    frac_w = a["pwfac"][1:, 1:mza - 1] / dzim[None, 1:mza - 1]
    assert np.abs(frac_w - exact_w[1:, 1:mza - 1]).max() < 3e-4
    iw = grid.arrays["v_iw"][1:] - 1
    exact_v = 0.5 * (1.0 / (1.0 + r[iw[:, 0]]) + 1.0 / (1.0 + r[iw[:, 1]]))
    dn = grid.arrays["dniv"][1:, None].astype(np.float64) * grid.arrays["zfacit"][None, :].astype(np.float64)
    frac_v = a["pvfac"][1:, 1:] / dn[:, 1:]
    assert np.abs(frac_v - exact_v[:, 1:]).max() < 3e-4
    # and they are not the total-water fractions' first-order mirror image
    assert np.abs(frac_w - exact_w[1:, 1:mza - 1]).max() < 0.1 * r.max()
