"""Pins of the CPU oracle (oracle/olam_oracle.cpp) against what the mathematics fixes.

The reference ships no golden vectors or unit tests for this path (its validation is
by idealised test cases only), so each oracle function is pinned by a property that a
plausible mistake (dropped term, wrong sign or index, transposed operand, wrong RK
stage weight) would violate:

  tridiffo (tridiag_lib.f90:10-36)       dense solve of random diagonally dominant systems
  update_scalar + RK driver              constant tracer stays constant (free stream);
    (scalar_transport_rk.F90:30-615)     FP64 tracer mass conserved to 1e-13; donor-cell
                                         (order 1) result equals an independent numpy
                                         implementation of the upwind scheme; linear
                                         vertical profile advected exactly by uniform w
                                         (pins grad_z_quadratic weights and signs);
                                         linearity in q
  adjoint (no reference counterpart)     dot-product test <L dq, l> = <dq, L^T l>, FP64
  implicit branch (:186-217, :347-375)   large vertical CFL: tracer mass conserved and
                                         constant tracer preserved through the solve
  check_cfls (:1476-1712)                closed-form CFL of a uniform vertical flux
  prog_wrtv_rk (prog_wrtv_rk.f90)        discrete hydrostatic rest state stays at rest;
                                         dry mass sum rho*volt conserved to 1e-14
  compute_mass_sums (check_nans.F90)     equals a numpy FP64 sum
  relabelling                            relabelled mesh gives bit-identical results

The option branches have their own pin files: order-3 reconstruction (test_order3.py,
test_mesh.py), monotone dynamics fluxes (test_monot_fluxes.py), Rayleigh friction
(test_rayleigh.py), eddy and horizontal diffusion (test_turbulence.py), moist pressure
factors (test_moist_factors.py), nudging (test_nudging.py) and all of them together
(test_full_config.py).
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from helpers import PermutedGrid, load_tracer_case, rel_linf, tracer_case
import oracle as O
from oracle import Oracle


def test_tridiffo_matches_dense_solve():
    rng = np.random.default_rng(3)
    for n in (2, 7, 59, 70):
        a = rng.standard_normal(n)
        c = rng.standard_normal(n)
        b = 2.5 + np.abs(a) + np.abs(c) + rng.random(n)
        d = rng.standard_normal(n)
        a[0] = 0.0
        c[-1] = 0.0
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        x = O.tridiffo(a, b, c, d)
        assert np.abs(x - np.linalg.solve(A, d)).max() < 1e-12


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_free_stream_and_mass(grid_factory, prec):
    grid = grid_factory(10, 16)
    w, st, tr = tracer_case(grid, ntr=3)
    o = Oracle(grid, abi.make_params(prec, num_scalar=3, dtsm=600.0))
    load_tracer_case(o, w, st, tr)
    st0 = st.copy()
    st0.arrays["rho"][:] = w["rho_old"].astype(np.float64)
    o.upload_state(st0)
    m0 = o.mass_sums()
    o.upload_state(st)
    o.scalar_transport_rk(w["rho_old"])
    m1 = o.mass_sums()
    q1 = o.download_scalar(1)[1:, 1:]
    assert np.abs(q1 - 1.0).max() == 0.0
    rel = np.abs(m1 - m0) / np.abs(m0)
    if prec == "f64":
        assert rel.max() < 1e-13
    else:
        assert rel[1:].max() < 1e-6


def _donor_cell_numpy(grid, w, q0, dtl):
    """Independent numpy implementation of the order-1 upwind RK3 step with no vertical
    flux: q_{s} = q0 + dt_s/(rho volt) sum_j dirv vm (q_up - q0)."""
    m = grid.mesh
    mza = grid.mza
    vm = w["vmasc"][1:].astype(np.float64)                  # (nv, mza)
    rho = w["rho"][1:].astype(np.float32).astype(np.float64)
    volt = grid.arrays["volt"][1:]
    q0 = q0[1:].astype(np.float64)
    q = q0.copy()
    for frac in (1 / 3, 1 / 2, 1.0):
        F = np.zeros_like(q)
        for j in range(7):
            act = j < m.npoly
            c = np.nonzero(act)[0]
            iv = m.w_iv[act, j]
            w1, w2 = m.v_iw[iv, 0], m.v_iw[iv, 1]
            up = np.where(vm[iv] < 0, q[w2], q[w1])
            F[c] += m.dirv[act, j][:, None] * vm[iv] * (up - q0[c])
        qn = q0.copy()
        qn[:, 1:] = q0[:, 1:] + frac * dtl * F[:, 1:] / volt[:, 1:] / rho[:, 1:]
        q = qn
    return q


def test_order1_matches_independent_donor_cell(grid_factory):
    grid = grid_factory(8, 12)
    w, st, tr = tracer_case(grid, ntr=1, wnoise=0.0)
    prm = abi.make_params("f64", num_scalar=1, dtsm=600.0, order=1)
    o = Oracle(grid, prm)
    st.arrays["wmasc"][:] = 0.0
    bell = cases.cosine_bell(grid)
    o.upload_state(st)
    o.upload_scalar(1, bell)
    o.scalar_transport_rk(w["rho_old"])
    got = o.download_scalar(1)[1:, 1:].astype(np.float64)
    ref = _donor_cell_numpy(grid, w, bell, 600.0)[:, 1:]
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()   # FP32 I/O of q


def test_linear_profile_advected_exactly_by_uniform_w(grid_factory):
    """Constant density, horizontally uniform q = a + b z, uniform vertical mass flux:
    every interior level moves by -w b dt exactly (face values of a linear profile are
    exact; RK3 integrates the constant tendency exactly)."""
    grid = grid_factory(6, 24)
    vg, cm = grid.vg, grid.cm
    mza = grid.mza
    rho0, wv, dt = 1.0, 0.5, 100.0
    st = abi.FortranState(grid)
    wm = np.zeros((mza + 1, grid.mesh.nw))
    for k in range(2, mza):
        wm[k] = rho0 * wv * cm["arw"][k]
    st.arrays["wmasc"][:] = grid.to_f(wm[1:].T.astype(np.float32))
    st.arrays["rho"][:] = rho0
    prm = abi.make_params("f64", num_scalar=1, dtsm=dt)
    o = Oracle(grid, prm)
    o.upload_state(st)
    zt = vg["zt"]
    q = 1.0 + 1e-4 * zt
    qf = np.repeat(q[1:][None, :], grid.mwa, 0).astype(np.float32)
    o.upload_scalar(1, qf)
    o.scalar_transport_rk(np.full((grid.mwa, mza), rho0, np.float32))
    got = o.download_scalar(1)[1:].astype(np.float64)
    expect = qf[1:].astype(np.float64) - wv * 1e-4 * dt
    # levels 5..mza-3 are beyond the reach of the two boundary faces after 3 stages
    err = np.abs(got[:, 4:mza - 3] - expect[:, 4:mza - 3]).max()
    # the shift is w b dt = 5e-3; the residual comes from FP32 storage of q and from the
    # deep-atmosphere face areas (zfacm^2 vs zfact^2, relative 1e-4 of the shift)
    assert err < 1e-5
    # and the boundary levels do change (the test is sensitive to the flux terms)
    assert np.abs(got[:, 1] - expect[:, 1]).max() > 1e-4


def test_vertical_upwinding_is_stable(grid_factory):
    """Many steps of vertical advection of a bump by uniform updraft (Courant 0.4) stay
    bounded: the face value is taken from the upwind cell (scalar_w_column picks
    scpb(k+1) when wmasc < 0, scpt(k) otherwise).  Choosing the downwind side makes the
    same integration grow by orders of magnitude."""
    grid = grid_factory(4, 40)
    vg, cm = grid.vg, grid.cm
    mza = grid.mza
    dz = vg["dzt"][2:].min()
    wv = 0.4 * dz / 100.0
    st = abi.FortranState(grid)
    wm = np.zeros((mza + 1, grid.mesh.nw))
    for k in range(2, mza):
        wm[k] = wv * cm["arw"][k]
    st.arrays["wmasc"][:] = grid.to_f(wm[1:].T.astype(np.float32))
    st.arrays["rho"][:] = 1.0
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=100.0))
    o.upload_state(st)
    zt = vg["zt"]
    q = np.exp(-((zt - 6000.0) / 1500.0) ** 2)
    o.upload_scalar(1, np.repeat(q[1:][None, :], grid.mwa, 0).astype(np.float32))
    for _ in range(60):
        o.scalar_transport_rk(np.ones((grid.mwa, mza), np.float32))
    out = o.download_scalar(1)[1:, 1:]
    assert np.isfinite(out).all() and np.abs(out).max() < 1.2 and out.max() > 0.5


def test_linearity_and_adjoint_dot_product(grid_factory):
    grid = grid_factory(8, 12)
    w, st, _ = tracer_case(grid, ntr=1)
    prm = abi.make_params("f64", num_scalar=3, dtsm=600.0)
    o = Oracle(grid, prm)
    o.upload_state(st)
    rng = np.random.default_rng(2)
    shape = (grid.mwa, grid.mza)
    q1 = rng.standard_normal(shape).astype(np.float32)
    q2 = rng.standard_normal(shape).astype(np.float32)
    q3 = (np.float32(2.0) * q1 - np.float32(0.5) * q2).astype(np.float32)
    for n, q in enumerate((q1, q2, q3), start=1):
        o.upload_scalar(n, q)
    o.scalar_transport_rk(w["rho_old"])
    L1, L2, L3 = (o.download_scalar(n)[1:, 1:].astype(np.float64) for n in (1, 2, 3))
    assert np.abs(L3 - (2.0 * L1 - 0.5 * L2)).max() < 1e-5 * np.abs(L3).max()
    lam = rng.standard_normal(shape).astype(np.float32)
    lam[:, 0] = 0
    o.upload_scalar_adj(1, lam)
    o.upload_scalar_adj(2, lam)
    o.upload_scalar_adj(3, lam)
    o.scalar_transport_rk_adj()
    LT = o.download_scalar_adj(1)[1:].astype(np.float64)
    a = float(np.sum(L1 * lam[1:, 1:]))
    b = float(np.sum(q1[1:].astype(np.float64) * LT))
    assert abs(a - b) / abs(a) < 1e-6


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_adjoint_dot_product_with_implicit_columns(grid_factory, mode):
    """Vertical Courant numbers far above 1 put every column through the implicit
    branch (tridif_prep/tridif_fini); level 1 keeps the explicit value there, which the
    adjoint must route to the explicit level 2 only after the implicit scp0 terms.  With
    mode > 0 the final stage is limited and its scales frozen (replayed forward)."""
    grid = grid_factory(8, 16)
    w, st, tr = tracer_case(grid, ntr=1, wnoise=40.0)
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=600.0, iscal_monot=mode))
    o.upload_state(st)
    o.upload_scalar(1, tr[0])
    o.scalar_transport_rk(w["rho_old"])
    assert o.cfl_report()[3] > 0.9 * (grid.mwa - 1)
    if mode:
        o.set_frozen_limiter(True)
    rng = np.random.default_rng(2)
    dq = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    dq[0] = 0
    lam[0] = 0
    o.upload_scalar(1, dq)
    o.scalar_transport_rk(w["rho_old"])
    Ldq = o.download_scalar(1)
    o.upload_scalar_adj(1, lam)
    o.scalar_transport_rk_adj()
    LT = o.download_scalar_adj(1)
    a = float(np.dot(Ldq[1:].ravel().astype(np.float64), lam[1:].ravel()))
    b = float(np.dot(dq[1:].ravel().astype(np.float64), LT[1:].ravel()))
    scale = float(np.linalg.norm(Ldq[1:].astype(np.float64)) * np.linalg.norm(lam[1:].astype(np.float64)))
    print(f"mode {mode}: mismatch / (|L dq| |lam|) {abs(a - b) / scale:.2e}")
    assert abs(a - b) / scale < 5e-9          # FP32 host arrays; a misplaced term gives 1e-6


def test_implicit_branch_conserves(grid_factory):
    grid = grid_factory(8, 16)
    w, st, tr = tracer_case(grid, ntr=2, wnoise=40.0)
    prm = abi.make_params("f64", num_scalar=2, dtsm=600.0)
    o = Oracle(grid, prm)
    load_tracer_case(o, w, st, tr)
    st0 = st.copy()
    st0.arrays["rho"][:] = w["rho_old"].astype(np.float64)
    o.upload_state(st0)
    m0 = o.mass_sums()
    o.upload_state(st)
    o.scalar_transport_rk(w["rho_old"])
    rep = o.cfl_report()
    assert rep[3] > 0 and rep[0] > 3.0
    m1 = o.mass_sums()
    assert (np.abs(m1 - m0) / np.abs(m0)).max() < 1e-12
    assert np.abs(o.download_scalar(1)[1:, 1:] - 1.0).max() < 1e-6


def test_cfl_closed_form(grid_factory):
    """Uniform upward mass flux w rho arw and no horizontal flow: total CFL at an
    interior level k is dt volti wm(k) / rho (outflow through the top face only)."""
    grid = grid_factory(6, 12)
    cm = grid.cm
    mza = grid.mza
    st = abi.FortranState(grid)
    wm = np.zeros((mza + 1, grid.mesh.nw))
    for k in range(2, mza):
        wm[k] = 0.7 * cm["arw"][k]
    st.arrays["wmasc"][:] = grid.to_f(wm[1:].T.astype(np.float32))
    st.arrays["rho"][:] = 1.0
    dt = 300.0
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=dt))
    o.upload_state(st)
    o.upload_scalar(1, np.ones((grid.mwa, mza), np.float32))
    o.scalar_transport_rk(np.ones((grid.mwa, mza), np.float32))
    rep = o.cfl_report()
    wmf = wm.astype(np.float32).astype(np.float64)
    expect = max((dt * wmf[k] / cm["volt"][k]).max() for k in range(2, mza))
    assert abs(rep[0] - expect) / expect < 1e-6
    assert rep[1] == 0.0


def test_rest_state_stays_at_rest():
    m = M.build_mesh(8)
    grid = abi.FortranGrid(m, M.vertical_grid(20))
    st = cases.jet_state(grid, u_jet=0.0, bump=0.0)
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=300.0))
    o.upload_state(st)
    o.tend0()
    dry0 = (st.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
    for _ in range(3):
        o.prog_wrtv_rk()
    out = abi.FortranState(grid)
    o.download_state(out)
    A = out.arrays
    assert np.abs(A["wc"][1:, 1:]).max() < 1e-6
    assert np.abs(A["vc"][1:, 1:]).max() < 1e-6
    assert np.abs(A["rho"][1:, 1:] / st.arrays["rho"][1:, 1:] - 1).max() < 1e-6      # residual vertical motion of the discrete balance
    assert np.abs(A["thil"][1:, 1:] - st.arrays["thil"][1:, 1:]).max() < 1e-4
    dry1 = (A["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
    assert abs(dry1 / dry0 - 1) < 1e-14


def test_unbalanced_state_moves_and_conserves_mass():
    """Sensitivity check for the test above: a 1 K bump does generate motion, and the
    dry mass is still conserved to roundoff."""
    m = M.build_mesh(8)
    grid = abi.FortranGrid(m, M.vertical_grid(20))
    st = cases.jet_state(grid, u_jet=0.0, bump=1.0)
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=300.0))
    o.upload_state(st)
    o.tend0()
    dry0 = (st.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
    for _ in range(3):
        o.prog_wrtv_rk()
    out = abi.FortranState(grid)
    o.download_state(out)
    assert np.abs(out.arrays["wc"][1:, 1:]).max() > 1e-4
    dry1 = (out.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
    assert abs(dry1 / dry0 - 1) < 1e-14


def test_mass_sums_match_numpy(grid_factory):
    grid = grid_factory(6, 10)
    w, st, tr = tracer_case(grid, ntr=2)
    o = Oracle(grid, abi.make_params("f64", num_scalar=2))
    load_tracer_case(o, w, st, tr)
    ms = o.mass_sums()
    mass = st.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]
    assert abs(ms[0] / mass.sum() - 1) < 1e-14
    assert abs(ms[2] / (mass * tr[1][1:, 1:].astype(np.float64)).sum() - 1) < 1e-13


def test_relabelled_mesh_bit_identical(grid_factory):
    base = grid_factory(6, 10)
    pg = PermutedGrid(base, seed=1)
    w, st, tr = tracer_case(base, ntr=2)
    jet = cases.jet_state(base)
    outs = []
    for grid, perm in ((base, None), (pg, pg)):
        fw = (lambda a, kind="w": a) if perm is None else perm.fwd
        bk = (lambda a, kind="w": a) if perm is None else perm.back
        o = Oracle(grid, abi.make_params("f32", num_scalar=2, dtsm=600.0))
        s2 = abi.FortranState(base)
        for name, kind, _ in abi.STATE_FIELDS:
            s2.arrays[name][...] = fw(jet.arrays[name], kind)
        s2.arrays["vmasc"][...] = fw(w["vmasc"], "v")
        s2.arrays["wmasc"][...] = fw(w["wmasc"], "w")
        o.upload_state(s2)
        for n, q in enumerate(tr, start=1):
            o.upload_scalar(n, fw(q))
        o.scalar_transport_rk(fw(w["rho_old"]))
        o.prog_wrtv_rk()
        res = abi.FortranState(base)
        o.download_state(res)
        outs.append([bk(o.download_scalar(2)), bk(res.arrays["thil"]), bk(res.arrays["vmc"], "v")])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_tracer_maxmin_matches_numpy(grid_factory):
    """tracer_maxmin (check_nans.F90:164-203): max and min of each tracer over levels
    lpw..mza of every real column; the dummy column and level 1 are left out (values
    planted there must not show), and a NaN is skipped."""
    grid = grid_factory(6, 10)
    rng = np.random.default_rng(4)
    o = Oracle(grid, abi.make_params("f32", num_scalar=3))
    qs = []
    for n in range(3):
        q = (rng.standard_normal((grid.mwa, grid.mza)) * (n + 1) - n).astype(np.float32)
        q[0, :] = 1e30          # dummy column
        q[:, 0] = -1e30         # level 1
        if n == 2:
            q[5, 4] = np.nan
        qs.append(q)
        o.upload_scalar(n + 1, q)
    mm = o.tracer_maxmin()
    for n, q in enumerate(qs):
        body = q[1:, 1:]
        assert mm[n, 0] == np.nanmax(body) and mm[n, 1] == np.nanmin(body)
