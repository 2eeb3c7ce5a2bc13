"""Oracle pins of the nudging terms of prog_wrtv_rk (prog_wrtv_rk.f90:347-355,605-631,
703-706,782-788), on the CPU.

* The nudging mass fluxes vmanud/wmanud are in flux form: the dry mass sum rho * volt is
  conserved to rounding in every mode.
* With nudflag > 1 and nud_preserve_total_mass the nudging fluxes enter the density and
  the long-step mass fluxes vmasc/wmasc alike, so a tracer's mass is conserved to
  rounding through a long step (a vmasc that leaves out vmanud breaks this by 1e-10).
  With nudflag == 1 the reference adds the vertical nudging flux to the density twice
  (:616 and :784) but to wmasc once; the oracle keeps that, so the tracer mass is not
  conserved there and the test does not ask it.
* rhot_nud (nud_preserve_total_mass off) changes the step through the explicit density
  only, leaving the mass conserved.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle


@pytest.fixture(scope="module")
def case():
    grid = abi.FortranGrid(M.build_mesh(6), M.vertical_grid(16))
    st = cases.jet_state(grid)
    rng = np.random.default_rng(1)
    st.arrays["thil"][1:] += rng.standard_normal(st.arrays["thil"][1:].shape).astype(np.float32)
    return grid, st, nudging_fields(grid, 1)


def nudging_fields(grid, seed):
    """Seeded nudging fields: wmanud zero at the surface and model-top W levels, vmanud
    on every face level, rhot_nud of 1e-6 of a typical density per second."""
    rng = np.random.default_rng(seed)
    wm = np.zeros((grid.mwa, grid.mza), np.float32)
    wm[1:, 1:-1] = 1e5 * rng.standard_normal((grid.mwa - 1, grid.mza - 2))
    vm = np.zeros((grid.mva, grid.mza), np.float32)
    vm[1:, 1:] = 1e5 * rng.standard_normal((grid.mva - 1, grid.mza - 1))
    rt = np.zeros((grid.mwa, grid.mza))
    rt[1:, 1:] = 1e-6 * rng.standard_normal((grid.mwa - 1, grid.mza - 1))
    return wm, vm, rt


def _long_step(grid, st, fields, nudflag, preserve):
    prm = abi.make_params("f64", num_scalar=1, dtsm=60.0, nudflag=nudflag, nud_preserve_total_mass=preserve)
    o = Oracle(grid, prm)
    o.upload_state(st)
    o.init_wrtv_rk()
    if nudflag:
        o.set_nudging(*fields)
    q0 = (1.0 + 0.5 * np.random.default_rng(5).standard_normal((grid.mwa, grid.mza))).astype(np.float32)
    o.upload_scalar(1, q0, None)
    m0 = o.mass_sums()
    o.tend0()
    o.prog_wrtv_rk()
    o.timeavg_momsc()
    o.scalar_transport_rk()
    m1 = o.mass_sums()
    out = abi.FortranState(grid)
    o.download_state(out)
    o.close()
    return (m1[0] - m0[0]) / m0[0], (m1[1] - m0[1]) / m0[1], out.arrays


@pytest.mark.parametrize("nudflag,preserve", [(1, 1), (2, 1), (1, 0)])
def test_dry_mass_conserved(case, nudflag, preserve):
    grid, st, f = case
    dm, _, _ = _long_step(grid, st, f, nudflag, preserve)
    assert abs(dm) < 2e-14


def test_tracer_mass_conserved_with_nudging_fluxes(case):
    grid, st, f = case
    _, dq, _ = _long_step(grid, st, f, 2, 1)
    assert abs(dq) < 5e-14


@pytest.mark.parametrize("nudflag,preserve", [(1, 1), (2, 1), (1, 0)])
def test_nudging_changes_the_step(case, nudflag, preserve):
    grid, st, f = case
    _, _, a = _long_step(grid, st, f, nudflag, preserve)
    _, _, b = _long_step(grid, st, f, 0, 1)
    assert not np.array_equal(a["press"], b["press"])
    if preserve:
        assert not np.array_equal(a["vmasc"], b["vmasc"])
        assert not np.array_equal(a["wmasc"], b["wmasc"])
    else:
        # rhot_nud enters only the explicit density: the mass fluxes of the first
        # stage are untouched, later stages see its effect through the pressure
        assert not np.array_equal(a["wmc"], b["wmc"])
