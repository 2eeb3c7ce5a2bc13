"""Oracle pins of Rayleigh friction near the model top (mem_rayf.f90:41-220,
prog_wrtv_rk.f90:46-52,219-303, vort_div_damp.f90:37-185), on the CPU.

* Profiles: the friction coefficients against their closed forms in the namelist
  values (a power law in height above zmin, zero below the first level above zmin;
  rayf on thil/vmc only with a horizontally homogeneous initialization).
* rayf_mix_top_vxe: a flux-form vertical diffusion, so it conserves each column's
  momentum, never adds kinetic energy (sum_k v * dv <= 0) and leaves a wind that is
  uniform in height unchanged.
* rayf on thil and vmc: zero where the field equals the reference, and every
  increment points toward the reference.
* rayfdiv: the second-order damping of the divergence lowers the divergence norm in
  the friction layer and leaves the levels below it alone.
* rayfw: W damping (b20 > 1 on the implicit diagonal) lowers |w| in the layer.

Tendency pins compare two oracle runs that differ only in the friction switch, with
the other damping off, through the oracle's setup_tendencies hook.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from oracle import Oracle

RF = {"rayf": (20000.0, 0.3, 1.0), "rayfw": (21000.0, 0.5, 1.0),
      "rayfdiv": (19000.0, 0.15, 1.0), "rayfmix": (22000.0, 0.25, 2.0)}


@pytest.fixture(scope="module")
def grid():
    g = abi.FortranGrid(M.build_mesh(8), M.vertical_grid(20))
    st = cases.jet_state(g)
    g.set_sounding(*cases.reference_sounding(g, st))
    return g


def _oracle(grid, st, prec="f64", **kw):
    kw.setdefault("divh_damp_level", 0)
    kw.setdefault("vort_damp_level", 0)
    o = Oracle(grid, abi.make_params(prec, dtsm=60.0, **kw))
    o.upload_state(st)
    o.init_wrtv_rk()
    return o


def _first_above(z, zmin, klo, khi):
    for k in range(klo, khi + 1):
        if z[k - 1] > zmin:
            return k
    return len(z) + 1


@pytest.mark.parametrize("initial", [1, 0])
def test_profiles_closed_form(grid, initial):
    st = cases.jet_state(grid)
    o = _oracle(grid, st, rayf=RF, initial=initial)
    p = o.rayf_profiles()
    o.close()
    mza, dt = grid.mza, 60.0
    zm, zt = grid.arrays["zm"].astype(np.float64), grid.arrays["zt"].astype(np.float64)
    top = zm[mza - 1]
    k = np.arange(1, mza + 1)

    def law(z, zmin, fact, expon, kbot, khi, norm_zmin=None):
        nz = zmin if norm_zmin is None else norm_zmin
        out = np.where((k >= kbot) & (k <= khi), fact * ((z - zmin) / (top - nz)) ** expon, 0.0)
        return np.where(np.isfinite(out), out, 0.0)

    zmin, fact, ex = RF["rayf"]
    kb = _first_above(zt, zmin, 2, mza) if initial == 1 else mza + 1
    assert p["krayf_bot"] == kb
    np.testing.assert_allclose(p["rayf_cof"], law(zt, zmin, fact / dt, ex, kb, mza), rtol=1e-6, atol=0)
    zmin, fact, ex = RF["rayfw"]
    kb = _first_above(zm, zmin, 2, mza - 1)
    assert p["krayfw_bot"] == kb
    cofw = law(zm, zmin, fact / dt, ex, kb, mza - 1)
    np.testing.assert_allclose(p["rayf_cofw"], cofw, rtol=1e-6, atol=0)
    np.testing.assert_allclose(p["b20"], 1.0 + dt * cofw, rtol=1e-6, atol=0)
    assert p["b20"][mza - 1] == 1.0 and p["b20"][0] == 1.0
    assert p["krayfdiv_bot"] == _first_above(zt, RF["rayfdiv"][0], 2, mza)
    zmin, fact, ex = RF["rayfmix"]
    kb = _first_above(zm, zmin, 2, mza - 1)
    assert p["krayfmix_bot"] == kb
    # the mixing profile is normalised by zm(mza) - rayfw_zmin (mem_rayf.f90:150-151)
    np.testing.assert_allclose(p["rayf_cofmix"], law(zm, zmin, 0.5 * fact, ex, kb, mza - 1, RF["rayfw"][0]),
                               rtol=1e-6, atol=0)


def test_profiles_off_by_default(grid):
    st = cases.jet_state(grid)
    o = _oracle(grid, st)
    p = o.rayf_profiles()
    o.close()
    for n in ("rayf_cof", "rayf_cofw", "rayf_cofmix"):
        assert not p[n].any(), n
    assert (p["b20"] == 1.0).all()
    assert p["krayf_bot"] == p["krayfw_bot"] == p["krayfdiv_bot"] == p["krayfmix_bot"] == grid.mza + 1


def _delta(grid, st, kind, field, **kw):
    rf = {kind: RF[kind]}
    if kind == "rayfmix":
        # its profile is normalised by zm(mza) - rayfw_zmin, which must not vanish
        rf["rayfw"] = (RF["rayfw"][0], 0.0, 1.0)
    on = _oracle(grid, st, rayf=rf, **kw)
    off = _oracle(grid, st, **kw)
    a, b = on.setup_tendencies(), off.setup_tendencies()
    kb = on.rayf_profiles()
    on.close(); off.close()
    return a[field] - b[field], kb


def test_rayfmix_conserves_and_dissipates(grid):
    st = cases.jet_state(grid)
    for fld, f in (("vxe", "vmxet_s"), ("vye", "vmyet_s"), ("vze", "vmzet_s")):
        d, kb = _delta(grid, st, "rayfmix", f)
        d = d[1:]                                    # real columns
        v = st.arrays[fld][1:].astype(np.float64)
        assert np.abs(d).max() > 0
        kbot = kb["krayfmix_bot"]
        assert not d[:, : kbot - 1].any()           # nothing below the layer
        colsum = d.sum(axis=1)
        assert np.all(np.abs(colsum) <= 1e-12 * np.abs(d).sum(axis=1) + 1e-300)
        assert np.all((v * d).sum(axis=1) <= 1e-12 * np.abs(v * d).sum(axis=1))
    # a wind uniform in height gets no mixing tendency at all
    st2 = st.copy()
    for fld in ("vxe", "vye", "vze"):
        a = st2.arrays[fld]
        a[:] = a[:, [5]]
    d, _ = _delta(grid, st2, "rayfmix", "vmxet_s")
    assert not d.any()


def test_rayf_relaxes_thil_and_vmc_toward_reference(grid):
    st = cases.jet_state(grid)
    th01d = grid.arrays["th01d"].astype(np.float64)
    rng = np.random.default_rng(5)
    st.arrays["thil"][1:] += rng.standard_normal(st.arrays["thil"][1:].shape).astype(np.float32)
    # one column equal to the sounding gets no thil friction
    st.arrays["thil"][7] = grid.arrays["th01d"]
    d, kb = _delta(grid, st, "rayf", "thilt_s")
    kbot = kb["krayf_bot"]
    assert kbot <= grid.mza and np.abs(d).max() > 0
    assert not d[:, : kbot - 1].any()
    assert not d[7].any()
    toward = d[1:, kbot - 1:] * (th01d[None, kbot - 1:] - st.arrays["thil"][1:, kbot - 1:])
    assert toward.min() >= 0.0
    # vc03d is the wind at init_wrtv_rk: no vmc friction until vc moves away from it
    on = _oracle(grid, st, rayf={"rayf": RF["rayf"]})
    off = _oracle(grid, st)
    assert np.array_equal(on.setup_tendencies()["vmt_s"], off.setup_tendencies()["vmt_s"])
    st2 = st.copy()
    st2.arrays["vc"][1:] += rng.standard_normal(st2.arrays["vc"][1:].shape).astype(np.float32)
    for o in (on, off):
        o.upload_state(st2)
    dv = on.setup_tendencies()["vmt_s"] - off.setup_tendencies()["vmt_s"]
    on.close(); off.close()
    assert not dv[:, : kbot - 1].any() and np.abs(dv).max() > 0
    pull = dv[1:, kbot - 1:] * (st.arrays["vc"][1:, kbot - 1:].astype(np.float64) -
                                st2.arrays["vc"][1:, kbot - 1:])
    assert pull.min() >= 0.0


def _divergence(grid, vmc):
    """sum over the faces of -arv * dirv * vmc / volt per cell (vort_div_damp.f90:140-153)."""
    A = grid.arrays
    mza = grid.mza
    iv = A["w_iv"][1:] - 1                            # (nw, 7) 1-based Fortran indices - 1
    npoly = A["w_npoly"][1:]
    dirv = A["w_dirv"][1:].astype(np.float64)
    flux = (A["arv"].astype(np.float64) * vmc)        # (mva, mza)
    div = np.zeros((grid.mwa - 1, mza))
    for j in range(7):
        act = j < npoly
        div[act] -= dirv[act, j, None] * flux[iv[act, j]]
    vol = A["volt"][1:].copy()
    vol[:, 0] = 1.0                                   # level 1 is below ground
    return div / vol


def test_rayfdiv_damps_divergence_in_the_layer(grid):
    st = cases.jet_state(grid)
    rng = np.random.default_rng(9)
    st.arrays["vmc"][1:] += 0.5 * rng.standard_normal(st.arrays["vmc"][1:].shape).astype(np.float32)
    d, kb = _delta(grid, st, "rayfdiv", "vmt_s")
    kbot = kb["krayfdiv_bot"]
    assert kbot <= grid.mza
    assert not d[:, : kbot - 1].any() and np.abs(d).max() > 0
    div0 = _divergence(grid, st.arrays["vmc"].astype(np.float64))
    ddiv = _divergence(grid, d)
    lay = slice(kbot - 1, grid.mza)
    assert (div0[:, lay] * ddiv[:, lay]).sum() < 0.0
    # the damped wind (one short step of the tendency) has less divergence in the layer
    dt = 60.0
    div1 = div0 + dt * ddiv
    for k in range(kbot - 1, grid.mza):
        assert np.linalg.norm(div1[:, k]) < np.linalg.norm(div0[:, k])


def test_rayfw_damps_w_in_the_layer(grid):
    st = cases.jet_state(grid)
    rng = np.random.default_rng(2)
    st.arrays["wmc"][1:, 1:-1] = 0.05 * rng.standard_normal(st.arrays["wmc"][1:, 1:-1].shape)
    res = {}
    for name, rf in (("on", {"rayfw": RF["rayfw"]}), ("off", None)):
        o = _oracle(grid, st, rayf=rf)
        kb = o.rayf_profiles()["krayfw_bot"]
        o.prog_wrtv_rk()
        out = abi.FortranState(grid)
        o.download_state(out)
        o.close()
        res[name] = out.arrays["wmc"][1:].astype(np.float64)
    lay = slice(kb - 1, grid.mza - 1)
    on, off = np.abs(res["on"][:, lay]).mean(axis=0), np.abs(res["off"][:, lay]).mean(axis=0)
    assert np.all(on < off)
