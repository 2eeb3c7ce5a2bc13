"""Pins of fluxes_wrtv_monot (prog_wrtv_rk.f90:1133-1554) in the oracle (CPU).

The monotone θ/wind fluxes replace fluxes_wrtv on the final acoustic RK stage when
ithil_monot / iwind_monot > 0.  The reference holds no expected outputs for them, so
they are pinned by what the construction guarantees and by sensitivity checks:
every face value is one number shared by the two cells, so Σ ρ θ_il V stays exact in
flux form (a renormalisation applied with different scales on the two sides of a face
breaks this); the limiter removes the overshoots the unlimited scheme makes at a sharp
θ_il step; and with the switches off the step is bit-identical to the default.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle


def _grid(nxp=8, nlev=20):
    m = M.build_mesh(nxp)
    return abi.FortranGrid(m, M.vertical_grid(nlev))


def _step_state(grid, amp=20.0):
    """The balanced jet with a sharp +amp K step of thil in a longitude sector at
    3-8 km (not rebalanced: it drives strong local motion)."""
    st = cases.jet_state(grid)
    lat, lon = cases._latlon_w(grid.mesh)
    zt = grid.vg["zt"]
    sec = ((lon % (2 * np.pi)) < 0.6 * np.pi) & (np.abs(lat) < 0.8)
    th = st.arrays["thil"]
    for k in range(2, grid.mza + 1):
        if 3000.0 < zt[k] < 8000.0:
            th[1:, k - 1] += np.where(sec, amp, 0.0).astype(np.float32)
    return st, sec


def _run(grid, st, nsteps, prec="f64", order=2, sums=None, **kw):
    o = Oracle(grid, abi.make_params(prec, num_scalar=1, dtsm=60.0, order=order, **kw))
    o.upload_state(st)
    o.tend0()
    if sums is not None:
        sums.append(o.rhothil_sum())
    for _ in range(nsteps):
        o.prog_wrtv_rk()
    if sums is not None:
        sums.append(o.rhothil_sum())
    out = abi.FortranState(grid)
    o.download_state(out)
    return out.arrays


@pytest.mark.parametrize("order", [2, 3])
def test_monotone_fluxes_conserve_rho_thil(order):
    grid = _grid()
    st, _ = _step_state(grid)
    sums = []
    _run(grid, st, 3, order=order, sums=sums, ithil_monot=1, iwind_monot=1)
    rel = abs(sums[1] / sums[0] - 1)
    print(f"order {order}: rho*thil change {rel:.3e}")
    assert rel < 1e-13


def test_switches_off_is_the_default_step():
    grid = _grid()
    st, _ = _step_state(grid)
    a = _run(grid, st, 2)
    b = _run(grid, st, 2, ithil_monot=0, iwind_monot=0)
    for f in ("thil", "rho", "vxe", "wmc"):
        assert np.array_equal(a[f], b[f])


def test_monotone_fluxes_limit_the_step_overshoot():
    grid = _grid(10, 20)
    st, sec = _step_state(grid)
    zt = grid.vg["zt"]
    lev = [k for k in range(2, grid.mza + 1) if 3000.0 < zt[k] < 8000.0]
    th0 = st.arrays["thil"][1:, :].astype(np.float64)
    hi = th0[:, [k - 1 for k in lev]].max(axis=0)
    res = {}
    for name, kw in (("unlimited", {}), ("monotone", dict(ithil_monot=1))):
        a = _run(grid, st, 6, **kw)
        th = a["thil"][1:, :].astype(np.float64)
        res[name] = float((th[:, [k - 1 for k in lev]] - hi).max())
    print(f"max thil above the initial layer maximum: {res}")
    assert res["unlimited"] > 1e-3
    assert res["monotone"] < 0.5 * res["unlimited"]
