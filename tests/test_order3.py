"""Pins of the order-3 horizontal reconstruction in the oracle (CPU).

grad_t2d_quadratic (grad_lib.f90:278-341) fits a quadratic to the neighbour values of
each W cell with the least-squares coefficients of mem_adv.f90:161-312, and the face
forms (e.g. scalar_transport_rk.F90:509-525) evaluate it at the V point of each side.
The reference holds no expected outputs for them; they are pinned by what the fit
guarantees: data sampled from a quadratic give back its coefficients, and the face
value is the quadratic at the V point (unbounded: s0 + P(dxps, dyps); bounded:
the polynomial with quadratic terms taken about their cell means).  A swapped term,
side or slot index fails these.  The order-3 transport step is then pinned by flux-form
mass conservation, free-stream preservation, the limiters' guarantees, and the
adjoint's dot-product test, as at order 2.
"""

import numpy as np
import pytest

from olam_amd import abi, cases
from helpers import load_tracer_case, tracer_case
from oracle import Oracle


def _isolated_cells(mesh, stride=1):
    """Cells whose closed neighbourhoods are pairwise disjoint (graph distance >= 3)."""
    taken = np.zeros(mesh.nw, bool)
    out = []
    for c in range(0, mesh.nw, stride):
        nb = mesh.w_iw[c, : mesh.npoly[c]]
        ring2 = set(nb.tolist()) | {c}
        for n in nb:
            ring2 |= set(mesh.w_iw[n, : mesh.npoly[n]].tolist())
        if taken[list(ring2)].any():
            continue
        out.append(c)
        taken[c] = True
        taken[nb] = True
    return np.array(out)


def _neighbour_xy(mesh):
    g = mesh.geo
    side = np.where(mesh.dirv < 0, 0, 1)
    f = np.float32
    xw = (g["dxps"].astype(f)[mesh.w_iv, side] * f(2)).astype(np.float64)
    yw = (g["dyps"].astype(f)[mesh.w_iv, side] * f(2)).astype(np.float64)
    return xw, yw


@pytest.mark.parametrize("bounded", [False, True])
def test_quadratic_coefficients_and_face_values(grid_factory, bounded):
    grid = grid_factory(10, 4)
    mesh = grid.mesh
    g = mesh.geo
    cells = _isolated_cells(mesh)
    assert len(cells) > 40
    xw, yw = _neighbour_xy(mesh)
    rng = np.random.default_rng(7)
    scale = np.array([1e-5, 1e-5, 1e-10, 1e-10, 1e-10])
    cc = rng.standard_normal((len(cells), 5)) * scale
    a0 = 1.0 + rng.random(len(cells))
    xx0, xy0, yy0 = (g[k].astype(np.float64) for k in ("xx0", "xy0", "yy0"))
    q = np.zeros((grid.mwa, grid.mza), np.float64)

    def poly(i, c, x, y):
        b = cc[i]
        if bounded:
            return (a0[i] + b[0] * x + b[1] * y + b[2] * (x * x - xx0[c]) + b[3] * (x * y - xy0[c])
                    + b[4] * (y * y - yy0[c]))
        return a0[i] + b[0] * x + b[1] * y + b[2] * x * x + b[3] * x * y + b[4] * y * y

    for i, c in enumerate(cells):
        np_ = mesh.npoly[c]
        # the cell mean is the value at the centre of the bounded form's quadratic part
        q[c + 1, :] = a0[i] if bounded else poly(i, c, 0.0, 0.0)
        for j in range(np_):
            q[mesh.w_iw[c, j] + 1, :] = poly(i, c, xw[c, j], yw[c, j])
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, order=3))
    coefs, faces = o.recon(q.astype(np.float32), 3, bounded_coef=bounded, bounded_off=bounded)
    worst_c = worst_f = 0.0
    for i, c in enumerate(cells):
        got = coefs[:, c + 1, 1]
        # the data carry the FP32 rounding of values ~2 (the field crosses as FP32)
        err = np.abs(got - cc[i]) / scale
        worst_c = max(worst_c, float(err.max()))
        assert np.all(err < 1e-4), (c, got, cc[i])
        for j in range(mesh.npoly[c]):
            iv = mesh.w_iv[c, j]
            side = 0 if mesh.dirv[c, j] < 0 else 1
            want = poly(i, c, float(np.float32(g["dxps"][iv, side])), float(np.float32(g["dyps"][iv, side])))
            got_f = faces[side, iv + 1, 1]
            # FP32 data: a few ulps of the largest neighbour value
            qmax = np.abs(q[np.r_[c, mesh.w_iw[c, : mesh.npoly[c]]] + 1, 1]).max()
            worst_f = max(worst_f, abs(got_f - want) / qmax)
            assert abs(got_f - want) < 1e-6 * qmax, (c, j, got_f, want)
    print(f"bounded={bounded}: {len(cells)} cells, coef err / scale {worst_c:.2e}, face err / max|q| {worst_f:.2e}")


def test_order3_wrong_offsets_are_detected(grid_factory):
    """Sensitivity: the unbounded fit evaluated with the bounded offsets (the dynamics'
    form, prog_wrtv_rk.f90:983-984,1057-1071) differs from the polynomial at the V point
    by the cell means of the quadratic terms, so the face test above separates them."""
    grid = grid_factory(10, 4)
    mesh = grid.mesh
    cells = _isolated_cells(mesh)
    xw, yw = _neighbour_xy(mesh)
    q = np.zeros((grid.mwa, grid.mza), np.float64)
    for c in cells:
        q[c + 1, :] = 1.0
        for j in range(mesh.npoly[c]):
            q[mesh.w_iw[c, j] + 1, :] = 1.0 + 1e-10 * xw[c, j] ** 2
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, order=3))
    _, fu = o.recon(q.astype(np.float32), 3, bounded_coef=False, bounded_off=False)
    _, fm = o.recon(q.astype(np.float32), 3, bounded_coef=False, bounded_off=True)
    c = cells[0]
    iv = mesh.w_iv[c, 0]
    side = 0 if mesh.dirv[c, 0] < 0 else 1
    diff = fu[side, iv + 1, 1] - fm[side, iv + 1, 1]
    assert abs(diff - 1e-10 * mesh.geo["xx0"][c]) < 1e-3 * abs(diff)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_order3_step_conserves_mass_and_free_stream(grid_factory, mode):
    grid = grid_factory(10, 12)
    w, st, tr = tracer_case(grid, ntr=3)
    o = Oracle(grid, abi.make_params("f64", num_scalar=3, dtsm=600.0, order=3, iscal_monot=mode))
    load_tracer_case(o, w, st, tr)
    st0 = st.copy()
    st0.arrays["rho"][:] = w["rho_old"].astype(np.float64)
    o.upload_state(st0)
    m0 = o.mass_sums()
    o.upload_state(st)
    o.scalar_transport_rk(w["rho_old"])
    m1 = o.mass_sums()
    assert np.abs(o.download_scalar(1)[1:, 1:] - 1.0).max() < 1e-12
    rel = np.abs(m1 - m0) / np.abs(m0)
    print(f"mode {mode}: mass changes {rel}")
    assert rel.max() < 1e-13


def _step_tracer(grid):
    lat, lon = cases._latlon_w(grid.mesh)
    q = ((lon % (2 * np.pi)) < 0.5 * np.pi).astype(np.float64)
    out = np.zeros((grid.mwa, grid.mza), np.float32)
    out[1:] = q[:, None]
    return out


def _run(grid, mode, order, nsteps, q, dt=4000.0):
    w, st, _ = tracer_case(grid, dt=dt, ntr=1)
    o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=dt, order=order, iscal_monot=mode))
    o.upload_state(st)
    o.upload_scalar(1, q)
    for _ in range(nsteps):
        o.scalar_transport_rk(w["rho_old"])
    return o, w


def test_order3_limiters_keep_their_guarantees(grid_factory):
    grid = grid_factory(10, 8)
    q = _step_tracer(grid)
    free = _run(grid, 0, 3, 4, q)[0].download_scalar(1)[1:, 1:]
    mono = _run(grid, 1, 3, 4, q)[0].download_scalar(1)[1:, 1:]
    pd = _run(grid, 2, 3, 4, q)[0].download_scalar(1)[1:, 1:]
    o2 = _run(grid, 0, 2, 4, q)[0].download_scalar(1)[1:, 1:]
    print(f"order 3 unlimited [{free.min():.3e}, {free.max():.3e}] monotone [{mono.min():.3e}, "
          f"{mono.max():.3e}] pd min {pd.min():.3e}")
    assert free.min() < -1e-3 or free.max() > 1 + 1e-3          # the scheme does overshoot
    assert np.abs(free - o2).max() > 1e-4                         # order 3 is not order 2
    assert mono.min() > -1e-6 and mono.max() < 1 + 1e-6
    assert pd.min() > -1e-10


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_order3_adjoint_dot_product(grid_factory, mode):
    grid = grid_factory(8, 10)
    q = _step_tracer(grid)
    o, w = _run(grid, mode, 3, 1, q)
    if mode:
        o.set_frozen_limiter(True)
    rng = np.random.default_rng(12)
    dq = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    dq[0] = 0
    lam[0] = 0
    lam[:, 0] = 0
    o.upload_scalar(1, dq)
    o.scalar_transport_rk(w["rho_old"])
    Ldq = o.download_scalar(1)
    o.upload_scalar_adj(1, lam)
    o.scalar_transport_rk_adj()
    LTlam = o.download_scalar_adj(1)
    a = float(np.dot(Ldq[1:].ravel().astype(np.float64), lam[1:].ravel()))
    b = float(np.dot(dq[1:].ravel().astype(np.float64), LTlam[1:].ravel()))
    scale = float(np.linalg.norm(Ldq[1:].astype(np.float64)) * np.linalg.norm(lam[1:].astype(np.float64)))
    print(f"order 3 mode {mode}: mismatch / (|L dq| |lam|) {abs(a - b) / scale:.2e}")
    assert abs(a - b) / scale < 5e-9


def test_order3_dynamics_conserve_mass_and_differ_from_order2():
    from olam_amd import mesh as M
    m = M.build_mesh(8)
    grid = abi.FortranGrid(m, M.vertical_grid(20))
    st = cases.jet_state(grid)
    outs = {}
    for order in (2, 3):
        o = Oracle(grid, abi.make_params("f64", num_scalar=1, dtsm=300.0, order=order))
        o.upload_state(st)
        o.tend0()
        dry0 = (st.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
        for _ in range(3):
            o.prog_wrtv_rk()
        out = abi.FortranState(grid)
        o.download_state(out)
        dry1 = (out.arrays["rho"][1:, 1:] * grid.arrays["volt"][1:, 1:]).sum()
        assert abs(dry1 / dry0 - 1) < 1e-14
        outs[order] = out.arrays
    d = np.abs(outs[3]["vxe"][1:, 1:] - outs[2]["vxe"][1:, 1:]).max()
    print(f"order 3 vs order 2 vxe max diff {d:.3e}")
    assert d > 1e-6
