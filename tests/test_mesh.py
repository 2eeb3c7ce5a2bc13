"""Mesh generator and metric pins (CPU).

The grid is an input of the hot path, built by a topology-equivalent restatement of
OLAM's MAKEGRID stage.  These checks pin it by closed forms and exactness properties:
counts N_W = 10 NXP^2 + 2, N_V = 30 NXP^2, N_M = 20 NXP^2 (icosahedron.f90:46-48),
Euler characteristic 2, 12 pentagons, dirv antisymmetry, area convergence to 4 pi a^2,
exactness of grad_t2d weights for linear fields (hex_grid.f90:757-806 and
mem_grid.f90:581-590), exact reconstruction of uniform tangent vectors by ecvec
(the DGELS constraint of hex_grid.f90:995-1003), exactly divergence-free discrete
streamfunction winds, and the vertical reconstruction weights (mem_adv.f90:58-117)
reproducing linear profiles.
"""

import math

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M


@pytest.mark.parametrize("nxp", [3, 8, 20])
def test_counts_and_topology(nxp):
    m = M.build_mesh(nxp)
    assert m.nw == 10 * nxp * nxp + 2
    assert m.nv == 30 * nxp * nxp
    assert m.nm == 20 * nxp * nxp
    assert m.nw - m.nv + m.nm == 2
    assert (m.npoly == 5).sum() == 12 and (m.npoly == 6).sum() == m.nw - 12
    # each V: two distinct W and two distinct M; dirv is +1 on iw2, -1 on iw1
    assert np.all(m.v_iw[:, 0] != m.v_iw[:, 1]) and np.all(m.v_im[:, 0] != m.v_im[:, 1])
    for j in range(6):
        act = j < m.npoly
        c = np.nonzero(act)[0]
        iv = m.w_iv[act, j]
        side2 = m.v_iw[iv, 1] == c
        assert np.all(m.dirv[act, j] == np.where(side2, 1.0, -1.0))
        # neighbour across iv(j) is the other W of that V
        other = np.where(side2, m.v_iw[iv, 0], m.v_iw[iv, 1])
        assert np.array_equal(other, m.w_iw[act, j])
    # every V appears in exactly two W lists with opposite dirv
    s = np.zeros(m.nv)
    cnt = np.zeros(m.nv, int)
    for j in range(6):
        act = j < m.npoly
        np.add.at(s, m.w_iv[act, j], m.dirv[act, j])
        np.add.at(cnt, m.w_iv[act, j], 1)
    assert np.all(cnt == 2) and np.all(s == 0)


def test_area_converges_to_sphere():
    errs = []
    for nxp in (10, 20, 40):
        m = M.build_mesh(nxp)
        a = m.geo["arw0"].sum()
        assert abs(a - m.geo["arm0"].sum()) / a < 1e-12      # same kites, two partitions
        errs.append(abs(a / (4 * math.pi * M.ERAD ** 2) - 1))
    assert errs[-1] < 1e-4
    assert errs[0] / errs[-1] > 10                             # converges (chord area error)


def test_gradient_weights_exact_for_linear_fields():
    m = M.build_mesh(12)
    g = m.geo
    xw = np.stack([g["xew"], g["yew"], g["zew"]], 1)
    rng = np.random.default_rng(0)
    for c in rng.choice(m.nw, 200, replace=False):
        lat, lon = math.radians(g["glatw"][c]), math.radians(g["glonw"][c])
        args = (math.cos(lat), math.sin(lat), math.cos(lon), math.sin(lon))
        if abs(xw[c, 0]) < 1e2 and abs(xw[c, 1]) < 1e2:
            args = (0.0, math.copysign(1.0, xw[c, 2]), 1.0, 0.0)
        d = xw[m.w_iw[c, : m.npoly[c]]] - xw[c]
        x, y = M._ps(d, *args)
        a, b = 3.0e-3, -1.7e-3
        f = a * x + b * y
        gx = np.dot(g["gxps_coef"][c, : m.npoly[c]], f)
        gy = np.dot(g["gyps_coef"][c, : m.npoly[c]], f)
        assert abs(gx - a) < 1e-10 * abs(a) * 1e3 and abs(gy - b) < 1e-10 * abs(b) * 1e3


def test_ecvec_satisfies_tangent_plane_identity():
    """hex_grid.f90:945-1066: ecvec_j = b_j n_j with n_j the unit tangential projection
    of the V normal and b the minimum-norm correction of the Perot weights such that
    sum_j b_j n_j n_j^T is the identity on the tangent plane."""
    m = M.build_mesh(10)
    g = m.geo
    vn = np.stack([g["vnx"], g["vny"], g["vnz"]], 1)
    wn = np.stack([g["wnx"], g["wny"], g["wnz"]], 1)
    ec = np.stack([g["ecvec_vx"], g["ecvec_vy"], g["ecvec_vz"]], 2)    # (nw, 7, 3)
    ok = g["ecvec_ok"]
    assert ok.mean() > 0.99
    for c in np.nonzero(ok)[0][::23]:
        np_ = m.npoly[c]
        v = vn[m.w_iv[c, :np_]]
        vps = v - (v @ wn[c])[:, None] * wn[c][None, :]
        vps /= np.linalg.norm(vps, axis=1, keepdims=True)
        Mt = np.einsum("ja,jb->ab", ec[c, :np_], vps)
        P = np.eye(3) - np.outer(wn[c], wn[c])
        assert np.abs(Mt - P).max() < 1e-9
        # weights stay within the DGELS acceptance window (0.03, 0.7)
        b = np.linalg.norm(ec[c, :np_], axis=1)
        assert b.min() > 0.03 and b.max() < 0.7


def test_streamfunction_winds_are_divergence_free():
    m = M.build_mesh(16)
    vg = M.vertical_grid(10)
    grid = abi.FortranGrid(m, vg)
    w = cases.transport_winds(grid, 600.0, noise=0.0, wnoise=0.0)
    vmasc = w["vmasc"][1:]                       # (nv, mza)
    div = np.zeros((m.nw, grid.mza))
    for j in range(7):
        act = j < m.npoly
        div[act] += m.dirv[act, j][:, None] * vmasc[m.w_iv[act, j]].astype(np.float64)
    scale = np.abs(vmasc).max()
    assert np.abs(div[:, 1:]).max() < 2e-6 * scale     # FP32 storage of vmasc only
    # and the density stays put
    assert np.abs(w["rho"][1:, 1:] / w["rho_old"][1:, 1:] - 1).max() < 1e-6


@pytest.mark.parametrize("nlev", [20, 60])
def test_vertical_weights_reproduce_linear_profiles(nlev):
    vg = M.vertical_grid(nlev)
    mza = vg["mza"]
    zt, zm = vg["zt"], vg["zm"]
    s = 3.0 + 0.002 * zt                                      # linear in z
    for name in ("a_w", "a_wu"):
        a = vg[name]
        for k in range(3, mza):
            ds_k, ds_km = s[k + 1] - s[k], s[k] - s[k - 1]
            bot = s[k] + ds_k * a[k, 0, 0] + ds_km * a[k, 1, 0]
            top = s[k] + ds_k * a[k, 0, 1] + ds_km * a[k, 1, 1]
            assert abs(bot - (3.0 + 0.002 * zm[k - 1])) < 1e-9
            assert abs(top - (3.0 + 0.002 * zm[k])) < 1e-9
    # the bounded weights integrate to the cell mean for a quadratic bump
    a = vg["a_w"]
    for k in range(3, mza):
        zb, ztp = vg["dzt_bot"][k], vg["dzt_top"][k]
        a0 = (ztp ** 3 + zb ** 3) / (3 * vg["dzt"][k])
        assert a0 > 0


def test_fortran_tables_roundtrip():
    m = M.build_mesh(5)
    grid = abi.FortranGrid(m, M.vertical_grid(8))
    A = grid.arrays
    assert A["w_iw"].shape == (m.nw + 1, 7) and A["v_iw"].shape == (m.nv + 1, 2)
    # Fortran indices are 1-based with dummy point 1
    assert A["w_iw"][1:, : 5].min() >= 2
    back = A["w_iw"][1:] - 2
    for j in range(5):
        assert np.array_equal(back[:, j], m.w_iw[:, j])


# ------------------------------------------------------------------ order-3 setup
def _neighbour_ps(m):
    """Neighbour positions (xw, yw) = 2 (dxps, dyps) of the shared face, FP32 then FP64."""
    g = m.geo
    act = np.arange(7)[None, :] < m.npoly[:, None]
    side = np.where(m.dirv < 0, 0, 1)
    f = np.float32
    xw = np.where(act, g["dxps"].astype(f)[m.w_iv, side] * f(2), 0).astype(np.float64)
    yw = np.where(act, g["dyps"].astype(f)[m.w_iv, side] * f(2), 0).astype(np.float64)
    return act, xw, yw


def test_cell_moments_match_polygon_boundary_integrals():
    """hex_quad's integrals of x^2, xy, y^2 over the cell against Green's theorem on
    the M polygon (a boundary sum independent of the fan decomposition)."""
    m = M.build_mesh(6)
    g = m.geo
    xw = np.stack([g["xew"], g["yew"], g["zew"]], 1)
    xm = np.stack([g["xem"], g["yem"], g["zem"]], 1)
    sl, cl = np.sin(np.radians(g["glatw"])), np.cos(np.radians(g["glatw"]))
    so, co = np.sin(np.radians(g["glonw"])), np.cos(np.radians(g["glonw"]))
    C = lambda a: np.repeat(a[:, None], 7, axis=1)
    px, py = M._ps(xm[m.w_im] - xw[:, None, :], C(cl), C(sl), C(co), C(so))
    act = np.arange(7)[None, :] < m.npoly[:, None]
    px, py = np.where(act, px, 0).astype(np.float32), np.where(act, py, 0).astype(np.float32)
    mom = M.cell_second_moments(px, py, m.npoly)
    for c in range(0, m.nw, 7):
        n = m.npoly[c]
        x, y = px[c, :n].astype(np.float64), py[c, :n].astype(np.float64)
        x1, y1 = np.roll(x, -1), np.roll(y, -1)
        cr = x * y1 - x1 * y
        sgn = np.sign(cr.sum())
        ixx = sgn * np.sum(cr * (x * x + x * x1 + x1 * x1)) / 12
        iyy = sgn * np.sum(cr * (y * y + y * y1 + y1 * y1)) / 12
        ixy = sgn * np.sum(cr * (x * y1 + 2 * x * y + 2 * x1 * y1 + x1 * y)) / 24
        ref = np.array([ixx, ixy, iyy])
        assert np.allclose(mom[c], ref, rtol=1e-12, atol=1e-12 * abs(ixx))
    # the fan's area is the cell area arw0 up to the projection's distortion
    area = np.zeros(m.nw)
    for j in range(7):
        on = j < m.npoly
        jp = np.where(on, (j - 1) % m.npoly, 0)
        ax, ay = px[np.arange(m.nw), jp].astype(float), py[np.arange(m.nw), jp].astype(float)
        area += np.where(on, 0.5 * np.abs(ax * py[:, j] - ay * px[:, j]), 0)
    assert np.abs(area / g["arw0"] - 1).max() < 1e-2          # ~0.5 % at 1000 km cells


def test_quadratic_lsq_coefficients_match_fp64_solve():
    """xy_hu / xy_h (the FP32 no-pivot LU of mem_adv.f90:190-312) against a pivoted FP64
    solve of the same normal equations."""
    m = M.build_mesh(12)
    g = m.geo
    act, xw, yw = _neighbour_ps(m)
    xx0, xy0, yy0 = (g[k].astype(np.float64)[:, None] for k in ("xx0", "xy0", "yy0"))
    for key, X in (("xy_hu", np.stack([xw, yw, xw * xw, xw * yw, yw * yw], 1)),
                   ("xy_h", np.stack([xw, yw, (xw * xw - xx0) * act, (xw * yw - xy0) * act,
                                      (yw * yw - yy0) * act], 1))):
        ref = np.linalg.solve(np.einsum("bin,bjn->bij", X, X), X)
        got = np.transpose(g[key], (0, 2, 1)).astype(np.float64)
        err = np.abs(got - ref).max(axis=2) / np.abs(ref).max(axis=2)
        print(key, "max rel err", err.max())
        assert err.max() < 5e-6
        assert np.all(g[key][~act] == 0)


def test_quadratic_fit_reproduces_quadratics():
    """Neighbour values of a quadratic s - s0 = c1 x + c2 y + c3 x^2 + c4 xy + c5 y^2
    give back (c1..c5) through xy_hu; with the quadratic terms taken about their cell
    means (x^2 - xx0, ...) they give it back through xy_h.  Face offsets are the
    quadratic terms at each side's V point (mem_adv.f90:322-354)."""
    m = M.build_mesh(10)
    g = m.geo
    act, xw, yw = _neighbour_ps(m)
    rng = np.random.default_rng(3)
    cc = rng.standard_normal(5) * np.array([1e-5, 1e-5, 1e-10, 1e-10, 1e-10])
    xx0, xy0, yy0 = (g[k].astype(np.float64)[:, None] for k in ("xx0", "xy0", "yy0"))
    d_u = cc[0] * xw + cc[1] * yw + cc[2] * xw * xw + cc[3] * xw * yw + cc[4] * yw * yw
    d_b = (cc[0] * xw + cc[1] * yw + cc[2] * (xw * xw - xx0) + cc[3] * (xw * yw - xy0)
           + cc[4] * (yw * yw - yy0))
    for key, d in (("xy_hu", d_u), ("xy_h", d_b)):
        rec = np.einsum("bnt,bn->bt", g[key].astype(np.float64), d * act)
        err = np.abs(rec / cc - 1).max()
        print(key, "reproduction rel err", err)
        assert err < 2e-5
    f = np.float32
    dx, dy = g["dxps"].astype(f), g["dyps"].astype(f)
    assert np.array_equal(g["xx0_vu"], dx * dx) and np.array_equal(g["xy0_vu"], dx * dy)
    assert np.array_equal(g["xx0_v"], dx * dx - g["xx0"][m.v_iw])
    assert np.array_equal(g["yy0_v"], dy * dy - g["yy0"][m.v_iw])
    # a wrong side (iw1 <-> iw2) would differ
    assert not np.array_equal(g["xx0_v"], dx * dx - g["xx0"][m.v_iw[:, ::-1]])
