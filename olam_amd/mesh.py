"""Global icosahedral-hexagonal (Voronoi) mesh with OLAM's table and metric layout.

This module produces the *inputs* of the RK3 dynamical core: the neighbour tables
``itab_w``/``itab_v``/``itab_m`` (mem_ijtabs.f90:45-114) and the grid metrics of
mem_grid.f90 and mem_adv.f90, for flat terrain (``lpw = lpv = lpm = 2``).  It is a
topology-equivalent restatement of OLAM's MAKEGRID stage, not a copy of it:

* W points are the vertices of an icosahedron whose 20 faces are subdivided NXP
  times (N_W = 10 NXP^2 + 2, icosahedron.f90:46-48), projected to the sphere.
* M points are the circumcentres of the resulting triangles (N_M = 20 NXP^2),
  V points are the triangle edges (N_V = 30 NXP^2).
* Metric definitions follow omodel/hex_grid.f90:418-1066 (chord lengths ``dnu``/
  ``dnv``, quarter-kite areas, polar-stereographic ``dxps``/``dyps``, the
  area-weighted triangle gradients ``gxps1/2`` scaled by ``farm``, and the
  DGELS minimum-norm correction of the Perot ``ecvec`` weights).
* Flat-terrain control volumes follow hex_grid.f90:1608-1976 with topography 0:
  ``arw = arw0*zfacm^2`` on interior W levels, ``arv = dnu*dzt*zfact``,
  ``volt = dzt*arw0*zfact^2``.
* Derived arrays follow mem_grid.f90:350-618 (``volti``, ``volwi``, ``volvi``,
  ``zwgt_*8``, ``gdz_wgt*8``, ``gxps_coef``) and mem_adv.f90:23-357 (``a_w``,
  ``a_wu``, ``xx_yy``, ``xx_yy_m``).

All arrays here are 0-based over real points (no dummy index); ``fortran_tables``
converts to the reference's 1-based convention with the dummy index 1, which is
what the C-ABI takes.  Vertical arrays are indexed by the Fortran level ``k``
(index 0 unused) so that ``k`` reads the same as in the reference.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# consts_coms.f90:15-21,106-107
ERAD = 6371.22e3
OMEGA = 7.29212e-5
CP = 1004.0
CV = 717.0
RDRY = CP - CV
P00 = 1.0e5
GRAV = 9.81
ROCV = RDRY / CV
CPOCV = CP / CV
ROCP = RDRY / CP
PC1 = (1.0 / P00) ** ROCV
ALPHA_DRY = PC1 * RDRY ** CPOCV  # mem_basic.f90:50


# ----------------------------------------------------------------------------
# icosahedron and its subdivision
# ----------------------------------------------------------------------------

def _icosahedron():
    """12 unit vertices (poles first) and 20 counter-clockwise faces."""
    lat = math.atan(0.5)
    v = [(0.0, 0.0, 1.0)]
    for i in range(5):
        lon = math.radians(72.0 * i)
        v.append((math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)))
    for i in range(5):
        lon = math.radians(72.0 * i + 36.0)
        v.append((math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), -math.sin(lat)))
    v.append((0.0, 0.0, -1.0))
    v = np.array(v)
    f = []
    for i in range(5):
        u0, u1 = 1 + i, 1 + (i + 1) % 5
        l0, l1 = 6 + i, 6 + (i + 1) % 5
        f.append((0, u0, u1))
        f.append((u0, l0, u1))
        f.append((u1, l0, l1))
        f.append((11, l1, l0))
    f = np.array(f, dtype=np.int64)
    # orient every face counter-clockwise seen from outside
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    s = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c)
    f[s < 0] = f[s < 0][:, [0, 2, 1]]
    return v, f


def _subdivide(nxp: int):
    """W-point unit vectors and CCW triangles of the NXP-subdivided icosahedron."""
    n = nxp
    v, faces = _icosahedron()
    edge_lut = -np.ones((12, 12), dtype=np.int64)
    ne = 0
    for fa in faces:
        for p, q in ((fa[0], fa[1]), (fa[1], fa[2]), (fa[2], fa[0])):
            p, q = min(p, q), max(p, q)
            if edge_lut[p, q] < 0:
                edge_lut[p, q] = ne
                ne += 1
    assert ne == 30
    nint = (n - 1) * (n - 2) // 2
    nw = 10 * n * n + 2
    base_edge = 12
    base_face = 12 + 30 * (n - 1)

    # local barycentric grid (i, j) of one face: P = ((n-i-j) A + i B + j C)/n
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    m = (ii + jj) <= n
    ii, jj = ii[m], jj[m]
    kk = n - ii - jj
    lut = -np.ones((n + 1, n + 1), dtype=np.int64)
    # interior local numbering (i >= 1, j >= 1, k >= 1)
    inter = (ii >= 1) & (jj >= 1) & (kk >= 1)
    loc = np.cumsum(inter) - 1

    xyz = np.zeros((nw, 3))
    gid_faces = []
    for fidx, (a, b, c) in enumerate(faces):
        gid = np.empty(ii.shape, dtype=np.int64)
        w = np.stack([kk, ii, jj], axis=1)      # weights of (a, b, c)
        verts = np.array([a, b, c])
        nz = (w > 0).sum(axis=1)
        # vertices
        sel = nz == 1
        gid[sel] = verts[np.argmax(w[sel], axis=1)]
        # edge points
        sel = nz == 2
        if sel.any():
            ws = w[sel]
            zero = np.argmin(ws, axis=1)
            pairs = np.array([[1, 2], [0, 2], [0, 1]])[zero]
            vu = verts[pairs[:, 0]]
            vv = verts[pairs[:, 1]]
            wu = ws[np.arange(len(ws)), pairs[:, 0]]
            swap = vu > vv
            lo = np.where(swap, vv, vu)
            hi = np.where(swap, vu, vv)
            wlo = np.where(swap, n - wu, wu)
            eid = edge_lut[lo, hi]
            gid[sel] = base_edge + eid * (n - 1) + (wlo - 1)
            pts = (wlo[:, None] * v[lo] + (n - wlo)[:, None] * v[hi])
            xyz[gid[sel]] = pts
        sel = nz == 3
        gid[sel] = base_face + fidx * nint + loc[sel]
        pts = (w[sel, 0:1] * v[a] + w[sel, 1:2] * v[b] + w[sel, 2:3] * v[c])
        xyz[gid[sel]] = pts
        lut[ii, jj] = gid
        # triangles of this face
        up_i, up_j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        mu = (up_i + up_j) <= n - 1
        ui, uj = up_i[mu], up_j[mu]
        t_up = np.stack([lut[ui, uj], lut[ui + 1, uj], lut[ui, uj + 1]], axis=1)
        md = (up_i + up_j) <= n - 2
        di, dj = up_i[md], up_j[md]
        t_dn = np.stack([lut[di + 1, dj], lut[di + 1, dj + 1], lut[di, dj + 1]], axis=1)
        gid_faces.append(t_up)
        gid_faces.append(t_dn)
    xyz[:12] = v
    xyz /= np.linalg.norm(xyz, axis=1, keepdims=True)
    tri = np.concatenate(gid_faces, axis=0)
    return xyz, tri


# ----------------------------------------------------------------------------
# vertical grid
# ----------------------------------------------------------------------------

def vertical_grid(nlev: int, ztop: float = 30000.0, ratio: float = 1.04):
    """Stretched levels 0..ztop with a geometric ratio (<= 1.15 as olam_grid.f90:563-572).

    Returns a dict of per-level arrays indexed by the Fortran level k (0 unused),
    k = 1..mza with mza = nlev + 1.  zm(1) = 0 is the ground, T level k lies between
    zm(k-1) and zm(k).
    """
    mza = nlev + 1
    r = ratio
    dz0 = ztop * (r - 1.0) / (r ** nlev - 1.0) if r != 1.0 else ztop / nlev
    zm = np.zeros(mza + 2)
    for k in range(2, mza + 1):
        zm[k] = zm[k - 1] + dz0 * r ** (k - 2)
    zm[mza] = ztop
    zm[0] = -(zm[2] - zm[1])  # zm(0): mirror below ground
    zm[mza + 1] = zm[mza] + (zm[mza] - zm[mza - 1])
    # T levels: geometric-mean placement like olam_grid.f90:520-525
    zt = np.zeros(mza + 2)
    for k in range(1, mza + 2):
        zt[k] = 0.5 * (zm[k - 1] + zm[k])
    zt[1] = zm[1] - 0.5 * (zm[2] - zm[1])
    arr = {}
    arr["mza"] = mza
    arr["zm"] = zm[: mza + 1].copy()
    arr["zt"] = zt[: mza + 1].copy()
    dzm = np.zeros(mza + 1)
    dzt = np.zeros(mza + 1)
    for k in range(1, mza + 1):
        dzm[k] = zt[k + 1] - zt[k]
        dzt[k] = zm[k] - zm[k - 1]
    arr["dzm"], arr["dzt"] = dzm, dzt
    arr["dzim"] = np.where(dzm > 0, 1.0 / np.where(dzm > 0, dzm, 1), 0.0)
    arr["dzit"] = np.where(dzt > 0, 1.0 / np.where(dzt > 0, dzt, 1), 0.0)
    zfacm = (ERAD + arr["zm"]) / ERAD
    zfact = (ERAD + arr["zt"]) / ERAD
    arr["zfacm"], arr["zfact"] = zfacm, zfact
    arr["zfacit"] = 1.0 / zfact
    arr["zfacm2"] = zfacm ** 2
    # mem_grid.f90:350-440 (alloc_gridz_other), gravity constant (olam_grid.f90:757)
    dzt_top = np.zeros(mza + 1)
    dzt_bot = np.zeros(mza + 1)
    dzt_top[1] = arr["zm"][1] - arr["zt"][1]
    dzt_bot[1] = dzt[1] - dzt_top[1]
    for k in range(2, mza + 1):
        dzt_top[k] = arr["zm"][k] - arr["zt"][k]
        dzt_bot[k] = arr["zt"][k] - arr["zm"][k - 1]
    arr["dzt_top"], arr["dzt_bot"] = dzt_top, dzt_bot
    zwgt_top8 = np.zeros(mza + 1)
    zwgt_bot8 = np.zeros(mza + 1)
    for k in range(1, mza):
        zwgt_top8[k] = dzt_bot[k + 1] * arr["dzim"][k]
        zwgt_bot8[k] = dzt_top[k] * arr["dzim"][k]
    zwgt_top8[mza] = zwgt_top8[mza - 1]
    zwgt_bot8[mza] = zwgt_bot8[mza - 1]
    arr["zwgt_top8"], arr["zwgt_bot8"] = zwgt_top8, zwgt_bot8
    arr["gdz_wgtp8"] = zwgt_top8 * GRAV
    arr["gdz_wgtm8"] = zwgt_bot8 * GRAV
    # vertical quadratic reconstruction weights, mem_adv.f90:58-117 (2x2 normal equations)
    a_w = np.zeros((mza + 1, 2, 2))
    a_wu = np.zeros((mza + 1, 2, 2))
    for k in range(2, mza + 1):
        for bounded in (True, False):
            a0 = (dzt_top[k] ** 3 + dzt_bot[k] ** 3) / (3.0 * dzt[k]) if bounded else 0.0
            z1 = np.array([-dzm[k - 1], dzm[k]])
            z2 = z1 * z1 - a0
            A = np.stack([z1, z2], axis=1)          # rows: the two neighbours
            # least-squares fit s(z) = g z + h (z^2 - a0) through ds(k-1)=-..., ds(k)
            # solution [g, h] = (A^T A)^-1 A^T [-ds(k-1), ds(k)]
            M = np.linalg.solve(A.T @ A, A.T)        # 2x2: maps (-ds(k-1), ds(k)) -> (g, h)
            # coefficient of ds(k) (column 1) and of ds(k-1) (column 0, with sign -1)
            g_dsk, h_dsk = M[0, 1], M[1, 1]
            g_dskm, h_dskm = -M[0, 0], -M[1, 0]
            zb, zt_ = dzt_bot[k], dzt_top[k]
            tgt = a_w if bounded else a_wu
            # value at bottom face (z = -dzt_bot) and top face (z = +dzt_top) minus cell value
            tgt[k, 0, 0] = h_dsk * (zb * zb - a0) - g_dsk * zb      # ds(k)   -> bottom
            tgt[k, 1, 0] = h_dskm * (zb * zb - a0) - g_dskm * zb    # ds(k-1) -> bottom
            tgt[k, 0, 1] = h_dsk * (zt_ * zt_ - a0) + g_dsk * zt_   # ds(k)   -> top
            tgt[k, 1, 1] = h_dskm * (zt_ * zt_ - a0) + g_dskm * zt_ # ds(k-1) -> top
    a_w[1] = a_w[2]
    a_wu[1] = a_wu[2]
    arr["a_w"], arr["a_wu"] = a_w, a_wu
    return arr


# ----------------------------------------------------------------------------
# the mesh
# ----------------------------------------------------------------------------

@dataclass
class Mesh:
    nxp: int
    nw: int
    nv: int
    nm: int
    # W
    npoly: np.ndarray = field(repr=False, default=None)
    w_iv: np.ndarray = field(repr=False, default=None)
    w_iw: np.ndarray = field(repr=False, default=None)
    w_im: np.ndarray = field(repr=False, default=None)
    dirv: np.ndarray = field(repr=False, default=None)
    # V
    v_iw: np.ndarray = field(repr=False, default=None)
    v_im: np.ndarray = field(repr=False, default=None)
    # M
    m_iw: np.ndarray = field(repr=False, default=None)
    m_iv: np.ndarray = field(repr=False, default=None)
    m_im: np.ndarray = field(repr=False, default=None)
    geo: dict = field(repr=False, default_factory=dict)
    w_canon: np.ndarray = field(repr=False, default=None)
    v_canon: np.ndarray = field(repr=False, default=None)
    m_canon: np.ndarray = field(repr=False, default=None)
    v_orient: np.ndarray = field(repr=False, default=None)


def build_mesh(nxp: int, order: str = "natural", seed: int = 0) -> Mesh:
    """Build the mesh topology and horizontal metrics.

    ``order`` = "natural" keeps the construction order (face-major), "shuffled"
    applies a seeded random permutation to W, V and M points (used to show that
    the device renumbering is exact and results are permutation invariant).
    """
    xyz, tri = _subdivide(nxp)
    nw = xyz.shape[0]
    nm = tri.shape[0]
    assert nw == 10 * nxp * nxp + 2 and nm == 20 * nxp * nxp

    w_canon = np.arange(nw)
    m_canon = np.arange(nm)
    if order == "shuffled":
        rng = np.random.default_rng(seed)
        pw = rng.permutation(nw)          # new index of old point
        inv = np.empty_like(pw)
        inv[pw] = np.arange(nw)
        xyz = xyz[inv]
        w_canon = inv
        tri = pw[tri]
        pm = rng.permutation(nm)
        tri = tri[pm]
        m_canon = pm

    # directed half-edges: triangle t is on the left of a->b
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    src = np.concatenate([a, b, c])
    dst = np.concatenate([b, c, a])
    third = np.concatenate([c, a, b])
    tid = np.tile(np.arange(nm), 3)
    key = src * nw + dst
    o = np.argsort(key, kind="stable")
    key_s, dst_s, third_s, tid_s, src_s = key[o], dst[o], third[o], tid[o], src[o]

    def lookup(s, d):
        idx = np.searchsorted(key_s, s * nw + d)
        assert np.all(key_s[idx] == s * nw + d)
        return idx

    # undirected edges (V points)
    lo = np.minimum(src, dst)
    hi = np.maximum(src, dst)
    ekey = np.unique(lo * nw + hi)
    nv = ekey.shape[0]
    assert nv == 30 * nxp * nxp
    v_iw = np.stack([ekey // nw, ekey % nw], axis=1)
    if order == "shuffled":
        rng2 = np.random.default_rng(seed + 1)
        pv = rng2.permutation(nv)
        v_iw = v_iw[pv]
        flip = rng2.random(nv) < 0.5
        v_iw[flip] = v_iw[flip][:, ::-1]
        ekey_sorted_order = np.argsort(np.minimum(v_iw[:, 0], v_iw[:, 1]) * nw
                                       + np.maximum(v_iw[:, 0], v_iw[:, 1]))
        ekey_sorted = (np.minimum(v_iw[:, 0], v_iw[:, 1]) * nw
                       + np.maximum(v_iw[:, 0], v_iw[:, 1]))[ekey_sorted_order]
    else:
        ekey_sorted_order = np.arange(nv)
        ekey_sorted = ekey

    def edge_id(p, q):
        k = np.minimum(p, q) * nw + np.maximum(p, q)
        idx = np.searchsorted(ekey_sorted, k)
        assert np.all(ekey_sorted[idx] == k)
        return ekey_sorted_order[idx]

    # M points of each V: im(2) left of iw1->iw2, im(1) left of iw2->iw1
    i12 = lookup(v_iw[:, 0], v_iw[:, 1])
    i21 = lookup(v_iw[:, 1], v_iw[:, 0])
    v_im = np.stack([tid_s[i21], tid_s[i12]], axis=1)

    # W neighbour fans, counter-clockwise; im(j) between iv(j) and iv(j+1)
    first = np.searchsorted(src_s, np.arange(nw))
    cnt = np.bincount(src, minlength=nw)
    npoly = cnt.astype(np.int64)
    assert set(np.unique(npoly)) <= {5, 6} and (npoly == 5).sum() == 12
    w_iw = np.zeros((nw, 7), dtype=np.int64)
    w_im = np.zeros((nw, 7), dtype=np.int64)
    cur = dst_s[first]
    for j in range(6):
        active = j < npoly
        idx = lookup(np.arange(nw)[active], cur[active])
        w_iw[active, j] = cur[active]
        w_im[active, j] = tid_s[idx]
        nxt = cur.copy()
        nxt[active] = third_s[idx]
        cur = nxt
    # closure check: after npoly steps we are back at the first neighbour
    assert np.all(cur == dst_s[first])
    w_iv = np.zeros((nw, 7), dtype=np.int64)
    dirv = np.zeros((nw, 7))
    cells = np.repeat(np.arange(nw)[:, None], 7, axis=1)
    act = np.arange(7)[None, :] < npoly[:, None]
    w_iv[act] = edge_id(cells[act], w_iw[act])
    dirv[act] = np.where(v_iw[w_iv[act], 1] == cells[act], 1.0, -1.0)

    # M tables: iw(1:3) CCW, iv(j) = edge (iw(j), iw(j+1)), im(j) across iv(j)
    m_iw = tri.copy()
    m_iv = np.stack([edge_id(tri[:, 0], tri[:, 1]), edge_id(tri[:, 1], tri[:, 2]),
                     edge_id(tri[:, 2], tri[:, 0])], axis=1)
    m_im = np.where(v_im[m_iv, 0] == np.arange(nm)[:, None], v_im[m_iv, 1], v_im[m_iv, 0])

    mesh = Mesh(nxp=nxp, nw=nw, nv=nv, nm=nm, npoly=npoly, w_iv=w_iv, w_iw=w_iw,
                w_im=w_im, dirv=dirv, v_iw=v_iw, v_im=v_im, m_iw=m_iw, m_iv=m_iv,
                m_im=m_im)
    # canonical (construction-order) ids, independent of any renumbering
    mesh.w_canon = w_canon
    mesh.m_canon = m_canon
    cw1, cw2 = w_canon[v_iw[:, 0]], w_canon[v_iw[:, 1]]
    ck = np.minimum(cw1, cw2) * nw + np.maximum(cw1, cw2)
    mesh.v_canon = np.argsort(np.argsort(ck, kind="stable"), kind="stable")
    mesh.v_orient = np.where(cw1 < cw2, 1.0, -1.0)
    _horizontal_metrics(mesh, xyz)
    return mesh


def _ps(d, coslat, sinlat, coslon, sinlon):
    """Polar-stereographic projection of displacement d (map_proj_ps.f90:562-608)."""
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    xq = -sinlon * dx + coslon * dy
    yq = coslat * dz - sinlat * (coslon * dx + sinlon * dy)
    zq = sinlat * dz + coslat * (coslon * dx + sinlon * dy)
    t = 2.0 * ERAD / np.maximum(1.0, 2.0 * ERAD + zq)
    return xq * t, yq * t


def _horizontal_metrics(mesh: Mesh, xyz_unit: np.ndarray):
    g = mesh.geo
    nw, nv, nm = mesh.nw, mesh.nv, mesh.nm
    xw = xyz_unit * ERAD
    # M points: circumcentres of the triangles (normal of the plane through 3 points)
    p0, p1, p2 = xyz_unit[mesh.m_iw[:, 0]], xyz_unit[mesh.m_iw[:, 1]], xyz_unit[mesh.m_iw[:, 2]]
    cm = np.cross(p1 - p0, p2 - p0)
    cm /= np.linalg.norm(cm, axis=1, keepdims=True)
    xm = cm * ERAD
    iw1, iw2 = mesh.v_iw[:, 0], mesh.v_iw[:, 1]
    im1, im2 = mesh.v_im[:, 0], mesh.v_im[:, 1]
    xv = 0.5 * (xw[iw1] + xw[iw2])
    xv *= (ERAD / np.linalg.norm(xv, axis=1))[:, None]
    # hex_grid.f90:543-575
    dnu = np.linalg.norm(xm[im1] - xm[im2], axis=1)
    un = (xm[im2] - xm[im1]) / dnu[:, None]
    dnv = np.linalg.norm(xw[iw1] - xw[iw2], axis=1)
    vn = (xw[iw2] - xw[iw1]) / dnv[:, None]
    # orientation: vn x un points outward
    assert np.all(np.einsum("ij,ij->i", np.cross(vn, un), xv) > 0)
    dvm1 = np.linalg.norm(xv - xm[im1], axis=1)
    dvm2 = np.linalg.norm(xv - xm[im2], axis=1)
    qk1 = 0.25 * dvm1 * dnv
    qk2 = 0.25 * dvm2 * dnv
    arm0 = (np.bincount(im1, weights=2.0 * qk1, minlength=nm)
            + np.bincount(im2, weights=2.0 * qk2, minlength=nm))
    arw0 = (np.bincount(iw1, weights=qk1 + qk2, minlength=nw)
            + np.bincount(iw2, weights=qk1 + qk2, minlength=nw))
    wn = xyz_unit.copy()
    raxis = np.sqrt(xw[:, 0] ** 2 + xw[:, 1] ** 2)
    glatw = np.degrees(np.arctan2(xw[:, 2], raxis))
    glonw = np.degrees(np.arctan2(xw[:, 1], xw[:, 0]))
    sinlat = xw[:, 2] / ERAD
    coslat = raxis / ERAD
    big = raxis >= 1.0e2
    sinlon = np.where(big, xw[:, 1] / np.maximum(raxis, 1e2), 0.0)
    coslon = np.where(big, xw[:, 0] / np.maximum(raxis, 1e2), 1.0)

    npoly = mesh.npoly
    J = np.arange(7)
    act = J[None, :] < npoly[:, None]
    jm1 = (J[None, :] - 1) % npoly[:, None]       # j1 = j2 - 1 (cyclic)
    jp1 = (J[None, :] + 1) % npoly[:, None]
    cells = np.repeat(np.arange(nw)[:, None], 7, axis=1)
    ivj = mesh.w_iv
    # kite pieces of each edge on the side of M sector j (between iv(j) and iv(j+1))
    # sector j contains M point im(j); edge iv(j) touches sector j-1 and j.
    mj = mesh.w_im
    kite_on = lambda iv, im: np.where(mesh.v_im[iv, 0] == im, qk1[iv], qk2[iv])
    farm = np.zeros((nw, 7))
    farv = np.zeros((nw, 7))
    arwi = 1.0 / arw0
    iv_next = np.take_along_axis(ivj, jp1, axis=1)
    farm[act] = (kite_on(ivj[act], mj[act]) + kite_on(iv_next[act], mj[act])) * np.repeat(arwi[:, None], 7, 1)[act]
    farv[act] = (qk1[ivj[act]] + qk2[ivj[act]]) * np.repeat(arwi[:, None], 7, 1)[act]

    # polar-stereographic coordinates of neighbours and V points in each W's plane
    C = lambda a: np.repeat(a[:, None], 7, axis=1)
    clat, slat, clon, slon = C(coslat), C(sinlat), C(coslon), C(sinlon)
    dwn = xw[mesh.w_iw] - xw[:, None, :]
    xwn, ywn = _ps(dwn, clat, slat, clon, slon)
    dvn = xv[ivj] - xw[:, None, :]
    xvn, yvn = _ps(dvn, clat, slat, clon, slon)
    # hex_grid.f90:757-806: triangle (iw, iw(j1), iw(j2)) gradient weights, stored at j1
    # here we store, at index j, the triangle (iw, iw(j), iw(j+1)) (M sector j).
    xa, ya = xwn, ywn
    xb = np.take_along_axis(xwn, jp1, 1)
    yb = np.take_along_axis(ywn, jp1, 1)
    det = xa * yb - xb * ya
    det = np.where(act, det, 1.0)
    vmagi = 1.0 / det
    gxps1 = np.where(act, yb * vmagi, 0.0) * farm
    gxps2 = np.where(act, -ya * vmagi, 0.0) * farm
    gyps1 = np.where(act, -xb * vmagi, 0.0) * farm
    gyps2 = np.where(act, xa * vmagi, 0.0) * farm
    # mem_grid.f90:581-590: neighbour n1 is point 1 of sector n1 and point 2 of sector n1-1
    gxps_coef = np.where(act, gxps1 + np.take_along_axis(gxps2, jm1, 1), 0.0)
    gyps_coef = np.where(act, gyps1 + np.take_along_axis(gyps2, jm1, 1), 0.0)
    # dxps/dyps: V point in the plane of iw1 (side 1) and iw2 (side 2)
    dxps = np.zeros((nv, 2))
    dyps = np.zeros((nv, 2))
    side = np.where(mesh.dirv < 0, 0, 1)
    dxps[ivj[act], side[act]] = xvn[act]
    dyps[ivj[act], side[act]] = yvn[act]

    # ecvec: hex_grid.f90:945-1066, DGELS minimum-norm correction of Perot weights
    vnj = vn[ivj]                                         # (nw,7,3)
    wnj = wn[:, None, :]
    vdotw = np.einsum("ijk,ijk->ij", vnj, np.broadcast_to(wnj, vnj.shape))
    vps = vnj - vdotw[..., None] * wnj
    vps /= np.where(act[..., None], np.linalg.norm(vps, axis=2, keepdims=True), 1.0)
    wx, wy, wz = wn[:, 0:1], wn[:, 1:2], wn[:, 2:3]
    pos = wz >= 0
    with np.errstate(divide="ignore", invalid="ignore"):
        fact = np.where(pos, (wy * vps[..., 0] - wx * vps[..., 1]) / (1.0 + wz),
                        (wy * vps[..., 0] - wx * vps[..., 1]) / (1.0 - wz))
    vrx = np.where(pos, vps[..., 0] * wz - vps[..., 2] * wx + wy * fact,
                   -vps[..., 0] * wz + vps[..., 2] * wx + wy * fact)
    vry = np.where(pos, vps[..., 1] * wz - vps[..., 2] * wy - wx * fact,
                   -vps[..., 1] * wz + vps[..., 2] * wy - wx * fact)
    vrx = np.where(act, vrx, 0.0)
    vry = np.where(act, vry, 0.0)
    fo = 2.0 * farv
    A = np.stack([vrx * vrx, vry * vry, vrx * vry], axis=1)     # (nw,3,7)
    rhs = np.stack([1.0 - np.sum(fo * A[:, 0], 1), 1.0 - np.sum(fo * A[:, 1], 1),
                    -np.sum(fo * A[:, 2], 1)], axis=1)
    AAt = np.einsum("nij,nkj->nik", A, A)
    y = np.linalg.solve(AAt, rhs[..., None])[..., 0]
    corr = np.einsum("nij,ni->nj", A, y)
    bco = fo + corr
    ok = np.all(np.where(act, (bco > 0.03) & (bco < 0.7), True), axis=1)
    bco = np.where(ok[:, None], bco, fo)
    ecvec = bco[..., None] * vps
    ecvec = np.where(act[..., None], ecvec, 0.0)

    g.update(dict(
        xew=xw[:, 0], yew=xw[:, 1], zew=xw[:, 2], xem=xm[:, 0], yem=xm[:, 1], zem=xm[:, 2],
        xev=xv[:, 0], yev=xv[:, 1], zev=xv[:, 2],
        wnx=wn[:, 0], wny=wn[:, 1], wnz=wn[:, 2],
        vnx=vn[:, 0], vny=vn[:, 1], vnz=vn[:, 2],
        unx=un[:, 0], uny=un[:, 1], unz=un[:, 2],
        dnu=dnu, dnv=dnv, dniu=1.0 / dnu, dniv=1.0 / dnv,
        arw0=arw0, arm0=arm0, glatw=glatw, glonw=glonw,
        farm=farm, farv=farv, gxps_coef=gxps_coef, gyps_coef=gyps_coef,
        gxps1=gxps1, gyps1=gyps1, gxps2=gxps2, gyps2=gyps2,
        dxps=dxps, dyps=dyps, ecvec_vx=ecvec[..., 0], ecvec_vy=ecvec[..., 1],
        ecvec_vz=ecvec[..., 2], ecvec_ok=ok,
    ))
    # mem_adv.f90:121-147 Laplacian weights
    g["xx_yy"] = np.where(act, (1.0 / dnv[ivj]) * dnu[ivj] * arwi[:, None], 0.0)
    g["xx_yy_m"] = dnv[mesh.m_iv] * (1.0 / dnu[mesh.m_iv]) / arm0[:, None]
    # M points of each W fan in the W's polar-stereographic plane (FP32 as hex_quad's xm)
    dmn = xm[mesh.w_im] - xw[:, None, :]
    xmn, ymn = _ps(dmn, clat, slat, clon, slon)
    g.update(_quadratic_fit(mesh, np.where(act, xmn, 0.0).astype(np.float32),
                            np.where(act, ymn, 0.0).astype(np.float32), arw0))


def cell_second_moments(xm, ym, npoly):
    """Integrals of x^2, xy, y^2 over each W cell (hex_quad, outils/quadrature.f90:10-81).

    The cell is the fan of triangles (W, M(j-1), M(j)) in the W's polar-stereographic
    plane, with the W point at the origin.  On a triangle with one vertex at the origin
    and the others at p, q the integrals of the quadratic monomials are exact:
    |det(p, q)|/12 * (px^2 + px qx + qx^2) for x^2, likewise for y^2, and
    |det(p, q)|/24 * (2 px py + px qy + qx py + 2 qx qy) for xy.  The reference
    evaluates the same integrals with a 2-point conical-product Gauss rule, which is
    exact for these integrands.  Arguments are (nw, 7) arrays; FP64 result (nw, 3).
    """
    xm = xm.astype(np.float64)
    ym = ym.astype(np.float64)
    nw = xm.shape[0]
    out = np.zeros((nw, 3))
    J = np.arange(7)
    for j in range(7):
        on = j < npoly
        jp = np.where(on, (j - 1) % np.maximum(npoly, 1), 0)
        px, py = xm[np.arange(nw), jp], ym[np.arange(nw), jp]
        qx, qy = xm[:, j], ym[:, j]
        ad = np.abs(px * qy - py * qx)
        out[:, 0] += np.where(on, ad / 12.0 * (px * px + px * qx + qx * qx), 0.0)
        out[:, 1] += np.where(on, ad / 24.0 * (2 * px * py + px * qy + qx * py + 2 * qx * qy), 0.0)
        out[:, 2] += np.where(on, ad / 12.0 * (py * py + py * qy + qy * qy), 0.0)
    del J
    return out


def _lsq5(X):
    """(X X^T)^-1 X for a batch of 5 x 7 design matrices X (FP32, padded columns 0),
    with the arithmetic of mem_adv.f90:190-312: normal matrix, LU without pivoting
    (ludcmp :396-446: unit lower L, upper U, diagonal stored as 1/U(i,i)), then
    forward substitution with L and back substitution with U applied to the rows of X,
    each row scaled by 1/U(i,i) first and the off-diagonal factors rescaled to match.
    """
    f = np.float32
    X = X.astype(f)
    nb = X.shape[0]
    A = np.einsum("bin,bjn->bij", X, X).astype(f)
    # Doolittle LU, column by column (ludcmp)
    for j in range(5):
        for i in range(j):
            s = A[:, i, j].copy()
            for k in range(i):
                s = s - A[:, i, k] * A[:, k, j]
            A[:, i, j] = s
        for i in range(j, 5):
            s = A[:, i, j].copy()
            for k in range(j):
                s = s - A[:, i, k] * A[:, k, j]
            A[:, i, j] = s
        if j != 4:
            dum = f(1.0) / A[:, j, j]
            for i in range(j + 1, 5):
                A[:, i, j] = A[:, i, j] * dum
    for i in range(5):
        A[:, i, i] = f(1.0) / A[:, i, i]
    h = np.transpose(A, (0, 2, 1)).copy()   # h[i, j]: i < j L(j, i); i > j U(j, i); 1/U(i, i)
    Y = X.copy()
    for i in range(4, -1, -1):
        Y[:, i] = Y[:, i] * h[:, i, i][:, None]
    # rescale column j = 5..1 of h (0-based 4..0)
    for j in range(4, -1, -1):
        d = h[:, j, j]
        for i in range(j + 1, 5):
            h[:, i, j] = h[:, i, j] * d
        for i in range(j):
            h[:, i, j] = h[:, i, j] * d
        for i in range(j + 1, 5):
            h[:, j, i] = h[:, j, i] / d
    # forward substitution: row i -= sum_{k < i} h(k, i) row k
    for i in range(1, 5):
        for k in range(i):
            Y[:, i] = Y[:, i] - h[:, k, i][:, None] * Y[:, k]
    # back substitution: row i -= sum_{k > i} h(k, i) row k, k from 5 down
    for i in range(3, -1, -1):
        for k in range(4, i, -1):
            Y[:, i] = Y[:, i] - h[:, k, i][:, None] * Y[:, k]
    return Y


def _quadratic_fit(mesh: Mesh, xmn, ymn, arw0):
    """Horizontal quadratic reconstruction coefficients of mem_adv.f90:161-354.

    For each W cell, the neighbours sit at (xw, yw) = 2 (dxps, dyps) of the shared V
    face in the cell's plane.  ``xy_hu`` (unbounded) fits s - s0 = c1 x + c2 y + c3 x^2
    + c4 xy + c5 y^2 by least squares over the neighbours; ``xy_h`` (bounded) fits the
    same with the quadratic terms taken relative to their cell means (x^2 - xx0, ...),
    so that the polynomial's mean over the cell is s0.  The face offsets ``xx0_v``
    (bounded) and ``xx0_vu`` (unbounded) are the quadratic terms at the V point of each
    side.  Layout: xy_h (nw, 7, 5) [cell, neighbour, term]; *_v (nv, 2) [face, side].
    """
    g = mesh.geo
    f = np.float32
    nw = mesh.nw
    npoly = mesh.npoly
    act = np.arange(7)[None, :] < npoly[:, None]
    dx32 = g["dxps"].astype(f)
    dy32 = g["dyps"].astype(f)
    side = np.where(mesh.dirv < 0, 0, 1)
    ivj = mesh.w_iv
    xw = np.where(act, dx32[ivj, side] * f(2.0), f(0.0)).astype(f)
    yw = np.where(act, dy32[ivj, side] * f(2.0), f(0.0)).astype(f)
    mom = cell_second_moments(xmn, ymn, npoly)
    arw0i = (1.0 / arw0).astype(f)
    xx0 = mom[:, 0].astype(f) * arw0i
    xy0 = mom[:, 1].astype(f) * arw0i
    yy0 = mom[:, 2].astype(f) * arw0i
    m = act.astype(f)
    Xb = np.stack([xw, yw, (xw * xw - xx0[:, None]) * m, (xw * yw - xy0[:, None]) * m,
                   (yw * yw - yy0[:, None]) * m], axis=1).astype(f)
    Xu = np.stack([xw, yw, xw * xw, xw * yw, yw * yw], axis=1).astype(f)
    xy_h = np.transpose(_lsq5(Xb), (0, 2, 1)) * m[..., None]
    xy_hu = np.transpose(_lsq5(Xu), (0, 2, 1)) * m[..., None]
    iw = mesh.v_iw
    out = dict(xy_h=xy_h.astype(f), xy_hu=xy_hu.astype(f), xx0=xx0, xy0=xy0, yy0=yy0)
    out["xx0_vu"] = (dx32 * dx32).astype(f)
    out["xy0_vu"] = (dx32 * dy32).astype(f)
    out["yy0_vu"] = (dy32 * dy32).astype(f)
    out["xx0_v"] = (out["xx0_vu"] - xx0[iw]).astype(f)
    out["xy0_v"] = (out["xy0_vu"] - xy0[iw]).astype(f)
    out["yy0_v"] = (out["yy0_vu"] - yy0[iw]).astype(f)
    return out


# ----------------------------------------------------------------------------
# 3-D metrics on flat terrain
# ----------------------------------------------------------------------------

def column_metrics(mesh: Mesh, vg: dict) -> dict:
    """arw, arv, volt (FP64), volti, volwi, volvi, pwfac, pvfac for lpw = lpv = 2.

    Arrays are (mza+1, n) indexed by Fortran level k (row 0 unused); row 1 is the
    below-ground dummy level.
    """
    mza = vg["mza"]
    g = mesh.geo
    K = mza + 1
    arw = np.zeros((K, mesh.nw))
    for k in range(2, mza):                       # interior W levels are open
        arw[k] = g["arw0"] * vg["zfacm2"][k]
    volt = np.zeros((K, mesh.nw))
    for k in range(2, mza + 1):
        volt[k] = vg["dzt"][k] * g["arw0"] * vg["zfact"][k] ** 2
    arv = np.zeros((K, mesh.nv))
    for k in range(2, mza + 1):
        arv[k] = g["dnu"] * vg["dzt"][k] * vg["zfact"][k]
    volti = np.zeros_like(volt)
    volti[2:] = 1.0 / volt[2:]
    volwi = np.zeros_like(volt)
    volwi[2:mza] = 1.0 / (volt[2:mza] + volt[3:mza + 1])
    iw1, iw2 = mesh.v_iw[:, 0], mesh.v_iw[:, 1]
    volvi = np.zeros((K, mesh.nv))
    volvi[2:] = 1.0 / (volt[2:, iw1] + volt[2:, iw2])
    pwfac = np.repeat(vg["dzim"][:, None], mesh.nw, axis=1)
    pvfac = vg["zfacit"][:, None] * g["dniv"][None, :]
    return dict(arw=arw, arv=arv, volt=volt, volti=volti, volwi=volwi, volvi=volvi,
                pwfac=pwfac, pvfac=pvfac)
