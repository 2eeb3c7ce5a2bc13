"""Pins of the oracle's momentum equations, damping terms, default reconstruction and
time stepping, against what the mathematics fixes (CPU only).

Each pin is paired with a mistake planted in a copy of the oracle's arithmetic (the
``plant`` test hook, olam_oracle.cpp oracle_plant) that the pin must catch by a wide
margin, so a pin that passes cannot be one that would pass anything.

* Steady rigid rotation (cases.rotating_state): an isothermal atmosphere in rigid
  rotation about the polar axis is an exact steady solution of the deep, compressible,
  non-hydrostatic equations that prog_wrtv_rk discretises, its pressure gradient
  balancing gravity, the full Coriolis force of prog_wrt_begs (prog_wrtv_rk.f90:675-681)
  and the centripetal acceleration carried by the face momentum fluxes (:635-661).  The
  wind change over a fixed time must fall at second order with the cell size.  A flipped
  Coriolis sign, a flipped pressure difference in prog_v_begs (:921-937) or flipped
  momentum face fluxes break the balance by orders of magnitude.
* Damping (vort_div_damp.f90:1-445): divh_damp leaves a streamfunction (non-divergent)
  wind alone and vort_damp a potential (irrotational) wind, to rounding; each removes
  energy from the field it acts on; the level-3 tendency scales as fact^3
  (facth :77, factb :107-110, fvorth :317-323).
* Order-2 horizontal reconstruction (grad_t2d grad_lib.f90:236-269 with the face formula
  of scalar_v_column scalar_transport_rk.F90:474-533 and fluxes_wrtv
  prog_wrtv_rk.f90:1045-1053): a field linear in a cell's plane has exact face values at
  the V points on that cell's side.
* Time stepping: scalar_transport_rk's RK3 (stage weights 1/3, 1/2, 1,
  scalar_transport_rk.F90:91-93) converges at third order in dt for a fixed
  non-divergent flow; prog_wrtv_rk converges at first order in dt, which is what its
  off-centred implicit weights fr = fp2 = fp3 = 0.7 (prog_wrtv_rk.f90:546-551) give.
* Williamson test 1 (cosine bell, one revolution of solid-body rotation): the error
  falls with the cell size.
"""

import math

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle

ZTOP = 10000.0      # a shallow domain keeps the vertical truncation below the horizontal one
NLEV = 20
DT, NSTEP = 100.0, 12

_grids = {}


def _grid(nxp, nlev=NLEV, ztop=ZTOP):
    key = (nxp, nlev, ztop)
    if key not in _grids:
        _grids[key] = abi.FortranGrid(M.build_mesh(nxp), M.vertical_grid(nlev, ztop=ztop))
    return _grids[key]


def _rotation_error(nxp, fcoriol, u0, plant=0, prec="f64"):
    """Relative rms and max of vc(t) - vc(0) after NSTEP acoustic steps from the steady
    rigid rotation."""
    grid = _grid(nxp)
    st = cases.rotating_state(grid, u0=u0, fcoriol=fcoriol)
    prm = abi.make_params(prec, num_scalar=1, dtsm=DT)
    prm.fcoriol = fcoriol
    o = Oracle(grid, prm)
    if plant:
        o.plant(plant)
    o.upload_state(st)
    o.tend0()
    for _ in range(NSTEP):
        o.prog_wrtv_rk()
    out = abi.FortranState(grid)
    o.download_state(out)
    o.close()
    dv = out.arrays["vc"][1:, 1:].astype(np.float64) - st.arrays["vc"][1:, 1:]
    return float(np.sqrt((dv ** 2).mean()) / u0), float(np.abs(dv).max() / u0)


@pytest.mark.parametrize("fcoriol,u0", [(2.0 * M.OMEGA, 20.0), (0.0, 60.0)], ids=["coriolis", "no-coriolis"])
def test_rigid_rotation_steady_at_second_order(fcoriol, u0):
    """Measured (FP64): rotating 6.1e-4, 1.6e-4, 4.0e-5 (rms, relative to u0) at NXP 8, 16,
    32; without Coriolis (centripetal balance only) 3.1e-4, 5.7e-5, 1.1e-5."""
    errs = [_rotation_error(n, fcoriol, u0)[0] for n in (8, 16, 32)]
    print("rms relative wind change", errs)
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0
    assert errs[2] < 6e-5


@pytest.mark.parametrize("plant,fcoriol,u0,factor", [
    (1, 2.0 * M.OMEGA, 20.0, 100.0),     # Coriolis sign (measured 566x)
    (2, 2.0 * M.OMEGA, 20.0, 100.0),     # prog_v_begs pressure difference (580x)
    (3, 0.0, 60.0, 30.0),                # momentum face fluxes, centripetal balance (104x)
], ids=["coriolis-sign", "pgf-sign", "momentum-flux-sign"])
def test_rigid_rotation_catches_planted_mistakes(plant, fcoriol, u0, factor):
    good = _rotation_error(16, fcoriol, u0)[0]
    bad = _rotation_error(16, fcoriol, u0, plant=plant)[0]
    print(f"plant {plant}: {bad:.3e} vs {good:.3e}")
    assert bad > factor * good


@pytest.mark.parametrize("prec", ["f32", "mixed"])
def test_rigid_rotation_reduced_precision(prec):
    """The FP32 and reference-precision arithmetic hold the balance as well as FP64."""
    e64 = _rotation_error(16, 2.0 * M.OMEGA, 20.0)[0]
    e = _rotation_error(16, 2.0 * M.OMEGA, 20.0, prec=prec)[0]
    assert e < 1.1 * e64


# ---------------------------------------------------------------- damping

def _rest_grid():
    return _grid(10, 8, 20000.0)


def _winds(grid, kind, amp=10.0):
    """Face velocities (Fortran (mva, mza)) of a smooth wind: 'psi' from an M-point
    streamfunction (discretely non-divergent), 'chi' from a W-point potential
    (discretely irrotational), scaled to about amp m/s."""
    mesh, g = grid.mesh, grid.mesh.geo
    if kind == "psi":
        x = np.stack([g["xem"], g["yem"], g["zem"]], 1) / M.ERAD
        f = x[:, 0] * x[:, 1] + 0.7 * x[:, 2] ** 3 - 0.4 * x[:, 1] * x[:, 2]
        i1, i2 = mesh.v_im[:, 0], mesh.v_im[:, 1]
        vn = (f[i1] - f[i2]) / g["dnu"]
    else:
        x = np.stack([g["xew"], g["yew"], g["zew"]], 1) / M.ERAD
        f = x[:, 0] * x[:, 2] - 0.6 * x[:, 1] ** 3 + 0.3 * x[:, 0] * x[:, 1]
        i1, i2 = mesh.v_iw[:, 0], mesh.v_iw[:, 1]
        vn = (f[i2] - f[i1]) / g["dnv"]
    vn = amp * vn / np.abs(vn).max()
    prof = 1.0 + 0.5 * np.sin(np.arange(grid.mza) * 0.7)
    vc = np.zeros((grid.mva, grid.mza), np.float32)
    vc[1:, 1:] = (vn[:, None] * prof[None, 1:]).astype(np.float32)
    return vc


def _damping_tendency(grid, vc, which, fact=None, level=3, plant=0):
    """vmt_short of the setup of prog_wrtv_rk with only one damping on, from a state at
    rest in the horizontal mean (rho uniform on each level) carrying the wind vc."""
    st = cases.jet_state(grid, u_jet=0.0, bump=0.0)
    rho = st.arrays["rho"]
    iw1, iw2 = grid.mesh.v_iw[:, 0] + 1, grid.mesh.v_iw[:, 1] + 1
    st.arrays["vc"][:] = vc
    vmc = np.zeros_like(vc)
    vmc[1:] = (vc[1:].astype(np.float64) * 0.5 * (rho[iw1] + rho[iw2])).astype(np.float32)
    st.arrays["vmc"][:] = vmc
    kw = dict(divh_damp_level=level if which == "div" else 0, vort_damp_level=level if which == "vort" else 0)
    if fact is not None:
        kw["divh_damp_fact" if which == "div" else "vort_damp_fact"] = fact
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=60.0, **kw))
    if plant:
        o.plant(plant)
    o.upload_state(st)
    t = o.setup_tendencies()["vmt_s"]
    o.close()
    return t[1:, 1:]


# The wind is FP32 (mem_basic.f90:8-20), so a discretely non-divergent (irrotational)
# wind has a divergence (vorticity) of FP32 rounding, ~1e-7 of the face terms; the
# level-3 damping applies two Laplacians to it, which amplify that grid-scale noise
# against the smooth field's own divergence by the square of the cells per wavelength.
# Measured ratios: level 1 1.8e-6 (div) and 6.4e-6 (vort); level 3 7.1e-5 and 1.3e-5.
@pytest.mark.parametrize("level,bound", [(1, 2e-5), (3, 3e-4)])
def test_divergence_damping_leaves_streamfunction_wind(level, bound):
    grid = _rest_grid()
    t_psi = _damping_tendency(grid, _winds(grid, "psi"), "div", level=level)
    t_chi = _damping_tendency(grid, _winds(grid, "chi"), "div", level=level)
    r = np.abs(t_psi).max() / np.abs(t_chi).max()
    print(f"divh_damp on psi / on chi: {r:.2e}")
    assert np.abs(t_chi).max() > 0 and r < bound


@pytest.mark.parametrize("level,bound", [(1, 2e-5), (3, 3e-4)])
def test_vorticity_damping_leaves_potential_wind(level, bound):
    grid = _rest_grid()
    t_chi = _damping_tendency(grid, _winds(grid, "chi"), "vort", level=level)
    t_psi = _damping_tendency(grid, _winds(grid, "psi"), "vort", level=level)
    r = np.abs(t_chi).max() / np.abs(t_psi).max()
    print(f"vort_damp on chi / on psi: {r:.2e}")
    assert np.abs(t_psi).max() > 0 and r < bound


def _energy_change(grid, vc, t):
    """Sum over faces of the tendency times the wind, weighted by the face volumes."""
    vv = grid.arrays["volvi"][1:, 1:]
    return float((t * vc[1:, 1:].astype(np.float64) / vv).sum())


@pytest.mark.parametrize("which,kind,plant", [("div", "chi", 5), ("vort", "psi", 7)])
def test_damping_removes_energy(which, kind, plant):
    """Each damping opposes the wind it acts on; a flipped coefficient sign (planted)
    adds energy instead."""
    grid = _rest_grid()
    vc = _winds(grid, kind)
    rng = np.random.default_rng(4)
    noisy = vc + np.float32(0.5) * rng.standard_normal(vc.shape).astype(np.float32)
    noisy[0] = 0
    noisy[:, 0] = 0
    for w in (vc, noisy):
        assert _energy_change(grid, w, _damping_tendency(grid, w, which)) < 0
        assert _energy_change(grid, w, _damping_tendency(grid, w, which, plant=plant)) > 0


@pytest.mark.parametrize("which", ["div", "vort"])
def test_damping_scales_as_fact_cubed(which):
    """Level 3: doubling the damping factor multiplies the tendency by 8 (facth, factb
    vort_div_damp.f90:77,107-110; fvorth :317-323); at level 2 by 4.  The planted
    exponent level-1 of divh_damp gives 4 at level 3."""
    grid = _rest_grid()
    kind = "chi" if which == "div" else "psi"
    vc = _winds(grid, kind)
    base = 0.1 if which == "div" else 0.075
    for level, expect in ((3, 8.0), (2, 4.0)):
        t1 = _damping_tendency(grid, vc, which, fact=base, level=level)
        t2 = _damping_tendency(grid, vc, which, fact=2 * base, level=level)
        m = np.abs(t1) > 1e-3 * np.abs(t1).max()
        ratio = t2[m] / t1[m]
        assert np.abs(ratio / expect - 1).max() < 1e-5, (level, ratio.min(), ratio.max())
    if which == "div":
        t1 = _damping_tendency(grid, vc, which, fact=base, plant=6)
        t2 = _damping_tendency(grid, vc, which, fact=2 * base, plant=6)
        m = np.abs(t1) > 1e-3 * np.abs(t1).max()
        assert np.abs(t2[m] / t1[m] / 8.0 - 1).min() > 0.4


# ---------------------------------------------------------------- order-2 reconstruction

def _disjoint_cells(mesh, n):
    taken = np.zeros(mesh.nw, bool)
    out = []
    rng = np.random.default_rng(5)
    for c in rng.permutation(mesh.nw):
        nb = mesh.w_iw[c, : mesh.npoly[c]]
        # closed neighbourhoods of the chosen cells must not touch
        ring = set(nb.tolist()) | {c}
        ring2 = set()
        for x in ring:
            ring2 |= set(mesh.w_iw[x, : mesh.npoly[x]].tolist())
        if taken[list(ring | ring2)].any():
            continue
        taken[list(ring)] = True
        out.append(c)
        if len(out) == n:
            break
    return out


@pytest.mark.parametrize("plant", [0, 4])
def test_order2_face_values_exact_for_linear_fields(plant):
    """q = f0 + a x + b y in a cell's polar-stereographic plane (map_proj_ps, the frame
    of gxps_coef and dxps): grad_t2d returns (a, b) and the face value on the cell's side
    of each of its faces is the linear field at the V point.  Planted: the other side's
    offsets (dxps(iw2) for iw1) are off by the full cell spacing times the gradient."""
    grid = _grid(12, 2)
    mesh, g = grid.mesh, grid.mesh.geo
    xw = np.stack([g["xew"], g["yew"], g["zew"]], 1)
    xv = np.stack([g["xev"], g["yev"], g["zev"]], 1)
    cells = _disjoint_cells(mesh, 60)
    q = np.full((grid.mwa, grid.mza), 0.25, np.float32)
    rng = np.random.default_rng(6)
    expect = {}
    for c in cells:
        lat, lon = math.radians(g["glatw"][c]), math.radians(g["glonw"][c])
        args = (math.cos(lat), math.sin(lat), math.cos(lon), math.sin(lon))
        if abs(xw[c, 0]) < 1e2 and abs(xw[c, 1]) < 1e2:
            args = (0.0, math.copysign(1.0, xw[c, 2]), 1.0, 0.0)
        a, b = rng.uniform(-2e-6, 2e-6, 2)
        f0 = rng.uniform(0.5, 1.5)
        nb = mesh.w_iw[c, : mesh.npoly[c]]
        for p in np.concatenate([[c], nb]):
            x, y = M._ps((xw[p] - xw[c])[None], *args)
            q[p + 1, :] = np.float32(f0 + a * x[0] + b * y[0])
        for j in range(mesh.npoly[c]):
            iv = mesh.w_iv[c, j]
            x, y = M._ps((xv[iv] - xw[c])[None], *args)
            side = 0 if mesh.v_iw[iv, 0] == c else 1
            expect[(iv, side)] = f0 + a * x[0] + b * y[0]
    o = Oracle(grid, abi.make_params("f64", num_scalar=1))
    if plant:
        o.plant(plant)
    _, faces = o.recon(q, 2)
    o.close()
    err = max(abs(faces[side, iv + 1, 1:] - v).max() for (iv, side), v in expect.items())
    print(f"plant {plant}: max face error {err:.2e}")
    if plant:
        assert err > 0.05
    else:
        assert err < 2e-6           # FP32 field values


# ---------------------------------------------------------------- time stepping

def _transport(grid, w, q0, T, n, plant=0, nrk=3):
    prm = abi.make_params("f64", num_scalar=1, dtsm=T / n, nrk_scal=nrk)
    o = Oracle(grid, prm)
    if plant:
        o.plant(plant)
    st = abi.FortranState(grid)
    st.arrays["vmasc"][:] = w["vmasc"]
    st.arrays["rho"][:] = w["rho_old"].astype(np.float64)
    o.upload_state(st)
    o.upload_scalar(1, q0)
    for _ in range(n):
        o.scalar_transport_rk(w["rho_old"])
    out = o.download_scalar(1)[1:, 1:].astype(np.float64)
    o.close()
    return out


def _self_convergence(runs):
    d = [np.abs(runs[i] - runs[i + 1]).max() for i in range(len(runs) - 1)]
    return [d[i] / d[i + 1] for i in range(len(d) - 1)]


@pytest.mark.parametrize("plant", [0, 9])
def test_tracer_rk3_third_order_in_time(plant):
    """Fixed non-divergent solid-body flow at constant density (a linear autonomous
    problem): successive halvings of dt shrink the change by 2^3.  Measured: 10.6, 9.9,
    8.4.  A wrong first-stage weight (planted 1/2) falls to second order (9.4, 5.3, 4.6)."""
    grid = _grid(16, 2, 30000.0)
    w = cases.transport_winds(grid, 1.0, u0=40.0, noise=0.0, wnoise=0.0)
    q0 = cases.cosine_bell(grid)
    runs = [_transport(grid, w, q0, 20000.0, n, plant) for n in (1, 2, 4, 8, 16)]
    r = _self_convergence(runs)
    print("ratios", r)
    if plant:
        assert r[-1] < 5.5
    else:
        assert 6.5 < r[-1] < 10.5 and r[-2] > 7.0


def test_dynamics_first_order_in_time():
    """prog_wrtv_rk over 6 s from an unbalanced state, dt = 2 s ... 0.125 s (vertical
    acoustic Courant numbers below 1, the asymptotic range): halving dt halves the
    change (measured ratios 1.92-2.07).  The off-centred weights fr = fp2 = fp3 = 0.7
    (prog_wrtv_rk.f90:546-551) of the vertically implicit acoustic terms and of the
    pressure extrapolation make the scheme first order; an error that did not vanish
    with dt (an inconsistent term) would not give these ratios."""
    grid = _grid(8, 10, 30000.0)
    st = cases.jet_state(grid, bump=2.0)
    T = 6.0
    runs = []
    for n in (3, 6, 12, 24, 48):
        prm = abi.make_params("f64", num_scalar=1, dtsm=T / n, divh_damp_level=0, vort_damp_level=0)
        o = Oracle(grid, prm)
        o.upload_state(st)
        o.tend0()
        for _ in range(n):
            o.prog_wrtv_rk()
        out = abi.FortranState(grid)
        o.download_state(out)
        o.close()
        runs.append(np.concatenate([out.arrays["wc"][1:, 1:].ravel(), out.arrays["vc"][1:, 1:].ravel()])
                    .astype(np.float64))
    r = _self_convergence(runs)
    print("ratios", r)
    assert all(1.7 < x < 2.3 for x in r)


def test_williamson1_cosine_bell_converges():
    """Cosine bell carried once around the equator by solid-body rotation at 40 m/s
    (Williamson et al. 1992, test 1, alpha = 0) at Courant ~0.8; the period includes the
    level's radius factor (arv, volt scale with zfact).  Relative l2 errors measured:
    0.62, 0.33, 0.11 at NXP 16, 32, 64 (0.033 at 128)."""
    errs = []
    u0 = 40.0
    for nxp in (16, 32, 64):
        grid = _grid(nxp, 1, 30000.0)
        w = cases.transport_winds(grid, 1.0, u0=u0, noise=0.0, wnoise=0.0)
        q0 = cases.cosine_bell(grid)
        period = 2.0 * math.pi * M.ERAD * grid.vg["zfact"][2] / u0
        q = _transport(grid, w, q0, period, 10 * nxp)
        qi = q0[1:, 1:].astype(np.float64)
        vol = grid.arrays["volt"][1:, 1:]
        errs.append(float(np.sqrt((vol * (q - qi) ** 2).sum() / (vol * qi ** 2).sum())))
    print("l2", errs)
    assert errs[0] / errs[1] > 1.6 and errs[1] / errs[2] > 2.5 and errs[2] < 0.15
