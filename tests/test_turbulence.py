"""Oracle pins of the eddy and horizontal diffusion in prog_wrtv_rk's setup loop
(solve_eddy_diff_heat / solve_eddy_diff_vxe, pbl_driver.f90:519-742; comp_horiz_k,
pbl_driver.f90:136-238; prog_wrtv_rk.f90:213-251), on the CPU.

* Vertical heat diffusion: the implicit fluxes equal a dense numpy solve of each
  column's backward-weighted diffusion system; the column's heat is conserved; a
  uniform column without tendency gets nothing.
* Vertical momentum diffusion: without surface drag each column's momentum is
  conserved; with drag a uniform wind loses momentum against itself.
* Horizontal diffusion: the face fluxes are antisymmetric, so every level's global sum
  of the volume-weighted tendency vanishes; with akmin > 0 the background diffusivity
  acts even where the PBL's vkm and vkh are zero, and without it zero fields do nothing.

Each pin compares oracle runs with and without the turbulence fields through the
oracle's setup_tendencies hook, with the other damping off, in FP64.  Fields that are
uniform along one direction switch off the other operator exactly: a field uniform in
each level gets no horizontal diffusion, a field uniform in each column (and without
tendency) gets no vertical heat diffusion.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle

DT = 60.0


@pytest.fixture(scope="module")
def grid():
    return abi.FortranGrid(M.build_mesh(6), M.vertical_grid(16))


def turbulence(grid, st, seed=1, drag=True):
    """Seeded PBL fields: vkh, vkm at W levels, tops khtop/kmtop, and a surface drag that
    removes about 10 % of the lowest level's momentum per acoustic step."""
    rng = np.random.default_rng(seed)
    mwa, mza = grid.mwa, grid.mza
    vkh = rng.uniform(5.0, 200.0, (mwa, mza)).astype(np.float32)
    vkm = rng.uniform(5.0, 200.0, (mwa, mza)).astype(np.float32)
    khtop = rng.integers(0, mza, mwa).astype(np.int32)
    kmtop = rng.integers(0, mza, mwa).astype(np.int32)
    akm = np.zeros(mwa, np.float32)
    if drag:
        A = grid.arrays
        dtomass = 1.1 * DT / A["volt"][1:, 1] / st.arrays["rho"][1:, 1]
        akm[1:] = (0.1 / (dtomass * A["dzit_bot"][1])).astype(np.float32)
    return vkh, vkm, khtop, kmtop, akm


def _oracle(grid, st, turb=None, akmin=0.0):
    o = Oracle(grid, abi.make_params("f64", dtsm=DT, divh_damp_level=0, vort_damp_level=0, akmin=akmin))
    o.upload_state(st)
    o.init_wrtv_rk()
    if turb is not None:
        o.set_turbulence(*turb)
    return o


def _delta(grid, st, turb, akmin=0.0):
    on, off = _oracle(grid, st, turb, akmin), _oracle(grid, st)
    a, b = on.setup_tendencies(), off.setup_tendencies()
    on.close(); off.close()
    return {k: a[k] - b[k] for k in a}


def _level_uniform(grid, st, fields, seed):
    rng = np.random.default_rng(seed)
    for f in fields:
        a = st.arrays[f]
        prof = a[1:].astype(np.float64).mean(axis=0) + rng.standard_normal(grid.mza)
        a[:] = prof[None, :].astype(np.float32)


def test_vertical_heat_diffusion_matches_dense_solve(grid):
    st = cases.jet_state(grid)
    _level_uniform(grid, st, ["thil"], 2)
    rng = np.random.default_rng(3)
    st.arrays["thilt"][1:] = 1e-3 * rng.standard_normal(st.arrays["thilt"][1:].shape).astype(np.float32)
    turb = turbulence(grid, st, drag=False)
    vkh, khtop = turb[0], turb[2]
    dv = _delta(grid, st, turb)["thilt_s"]
    A = grid.arrays
    rho = st.arrays["rho"]
    thil = st.arrays["thil"].astype(np.float64)
    checked = 0
    for iw in range(1, grid.mwa):
        kz = int(khtop[iw])
        if kz < 2:
            assert not dv[iw].any()
            continue
        ci = np.arange(2, kz + 2) - 1                      # T levels ka..kmax (column index)
        dtom = DT / A["volt"][iw, ci] / rho[iw, ci]        # the FP64 oracle rebuilds volti from volt
        dtomass = 1.1 * dtom
        varp = thil[iw, ci] + dtom * st.arrays["thilt"][iw, ci] * A["volt"][iw, ci]
        wi = np.arange(2, kz + 1) - 1                      # W levels ka..khtop
        akodz = A["arw"][iw, wi].astype(np.float64) * vkh[iw, wi] * A["dzim"][wi]
        n = len(wi)
        # flux x(k) = akodz(k) [(varp(k) - dtomass(k) (x(k) - x(k-1)))
        #                       - (varp(k+1) - dtomass(k+1) (x(k+1) - x(k)))]
        Mx = np.zeros((n, n))
        for i in range(n):
            Mx[i, i] = 1.0 + akodz[i] * (dtomass[i] + dtomass[i + 1])
            if i > 0:
                Mx[i, i - 1] = -akodz[i] * dtomass[i]
            if i < n - 1:
                Mx[i, i + 1] = -akodz[i] * dtomass[i + 1]
        x = np.linalg.solve(Mx, akodz * (varp[:-1] - varp[1:]))
        fl = np.concatenate([[0.0], x, [0.0]])
        expect = fl[:-1] - fl[1:]
        got = dv[iw, ci]
        assert np.abs(got - expect).max() <= 1e-9 * np.abs(expect).max() + 1e-300
        assert abs(got.sum()) <= 1e-12 * np.abs(got).sum()
        assert not dv[iw, kz + 1:].any() and dv[iw, 0] == 0
        checked += 1
    assert checked > 100


def test_vertical_heat_diffusion_uniform_column(grid):
    st = cases.jet_state(grid)
    st.arrays["thil"][:] = 300.0
    d = _delta(grid, st, turbulence(grid, st, drag=False))["thilt_s"]
    assert not d.any()


def test_vertical_momentum_diffusion_conserves(grid):
    st = cases.jet_state(grid)
    _level_uniform(grid, st, ["vxe", "vye", "vze"], 4)
    d = _delta(grid, st, turbulence(grid, st, drag=False))
    for f in ("vmxet_s", "vmyet_s", "vmzet_s"):
        dv = d[f][1:]
        assert np.abs(dv).max() > 0
        assert np.all(np.abs(dv.sum(axis=1)) <= 1e-12 * np.abs(dv).sum(axis=1) + 1e-300)


def test_surface_drag_opposes_the_wind(grid):
    st = cases.jet_state(grid)
    for f, u in (("vxe", 7.0), ("vye", -3.0), ("vze", 1.5)):
        st.arrays[f][:] = u
    d = _delta(grid, st, turbulence(grid, st, drag=True))
    for f, u in (("vmxet_s", 7.0), ("vmyet_s", -3.0), ("vmzet_s", 1.5)):
        colsum = d[f][1:].sum(axis=1)
        assert np.all(colsum * u < 0.0)
    # without drag a uniform wind is left alone (up to the rounding of the column solve)
    d0 = _delta(grid, st, turbulence(grid, st, drag=False))
    assert np.abs(d0["vmxet_s"]).max() <= 1e-9 * np.abs(d["vmxet_s"]).max()


def test_horizontal_diffusion_conserves(grid):
    st = cases.jet_state(grid)
    rng = np.random.default_rng(7)
    for f in ("thil", "vxe", "vye", "vze"):
        a = st.arrays[f]
        a[1:] = (a[1:, [5]] + rng.standard_normal((grid.mwa - 1, 1))).astype(np.float32)
    d = _delta(grid, st, turbulence(grid, st, drag=False))
    for f in ("thilt_s", "vmxet_s", "vmyet_s", "vmzet_s"):
        x = d[f][1:]
        assert np.abs(x).max() > 0, f
        assert np.all(np.abs(x.sum(axis=0)) <= 1e-10 * np.abs(x).sum(axis=0) + 1e-300), f


def test_akmin_background_diffusion(grid):
    st = cases.jet_state(grid)
    rng = np.random.default_rng(8)
    st.arrays["thil"][1:] += rng.standard_normal(st.arrays["thil"][1:].shape).astype(np.float32)
    z = np.zeros((grid.mwa, grid.mza), np.float32)
    zi = np.zeros(grid.mwa, np.int32)
    zero = (z, z, zi, zi, np.zeros(grid.mwa, np.float32))
    d = _delta(grid, st, zero, akmin=0.5)["thilt_s"][1:]
    assert np.abs(d[:, 1:]).min() > 0
    assert np.all(np.abs(d.sum(axis=0)) <= 1e-10 * np.abs(d).sum(axis=0) + 1e-300)
    assert not _delta(grid, st, zero, akmin=0.0)["thilt_s"].any()
