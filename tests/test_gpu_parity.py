"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on identical inputs.

Tolerances (relative to the field's max |value|, levels 2..mza):
  * FP64 mode: 1e-11 after one step (differences come only from FMA contraction,
    reciprocal-vs-division and the PCR-vs-Thomas column solve); 1e-9 for the vertical
    momentum fields, whose column solve amplifies rounding by ~1e6.
  * FP32 / MIXED modes, tracers: 2e-6 per RK3 step; dynamics: 1e-5 per acoustic step
    for rho/press/thil/velocities, 1e-4 for the vertical momentum fields whose max is
    small (wmc, wc, wmasc).  These are a few FP32 ulps of accumulated rounding; the
    column solve is FP64 on the GPU and FP64 in the oracle's F32/MIXED modes.
"""

import numpy as np
import pytest

from olam_amd import abi, cases
from helpers import DYN_FIELDS, dyn_fields, load_tracer_case, rel_l2, rel_linf, tracer_case
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_TR = {"f64": 1e-11, "f32": 2e-6, "mixed": 2e-6}


def _gpu(grid, prm):
    from olam_amd.gpu import Context
    return Context(grid, prm)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("nrk", [3, 2])
def test_tracer_step_parity(grid_factory, prec, nrk):
    grid = grid_factory(12, 20)
    w, st, tr = tracer_case(grid, ntr=4)
    prm = abi.make_params(prec, num_scalar=4, nrk_scal=nrk, dtsm=600.0)
    g, o = _gpu(grid, prm), Oracle(grid, prm)
    for be in (g, o):
        load_tracer_case(be, w, st, tr)
        be.scalar_transport_rk(w["rho_old"])
    for n in range(1, 5):
        e = rel_linf(g.download_scalar(n), o.download_scalar(n))
        print(f"{prec} nrk={nrk} tracer {n}: rel Linf {e:.3e}")
        assert e < TOL_TR[prec], (n, e)
    # a constant tracer stays exactly constant on the GPU (free-stream preservation)
    q1 = g.download_scalar(1)[1:, 1:]
    assert np.abs(q1 - 1.0).max() < (1e-12 if prec == "f64" else 2e-6)


def test_tracer_c1_config(grid_factory):
    """The CPU-runnable configuration: NXP=40, 30 levels, cosine bell, one RK3 step, FP64."""
    grid = grid_factory(40, 30)
    w, st, tr = tracer_case(grid, dt=300.0, ntr=2)
    prm = abi.make_params("f64", num_scalar=2, dtsm=300.0)
    g, o = _gpu(grid, prm), Oracle(grid, prm)
    for be in (g, o):
        load_tracer_case(be, w, st, tr)
        be.scalar_transport_rk(w["rho_old"])
    e = rel_linf(g.download_scalar(2), o.download_scalar(2))
    print(f"C1 cosine bell rel Linf {e:.3e}")
    assert e < 1e-11


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tracer_mass_conservation(grid_factory, prec):
    grid = grid_factory(16, 20)
    w, st, tr = tracer_case(grid, ntr=3)
    prm = abi.make_params(prec, num_scalar=3, dtsm=600.0)
    g = _gpu(grid, prm)
    load_tracer_case(g, w, st, tr)
    st0 = st.copy()
    st0.arrays["rho"][:] = w["rho_old"].astype(np.float64)
    g.upload_state(st0, only=["rho"])
    m0 = g.mass_sums()
    g.upload_state(st, only=["rho"])
    g.scalar_transport_rk(w["rho_old"])
    m1 = g.mass_sums()
    rel = np.abs((m1 - m0) / m0)
    print(prec, "mass rel change", rel)
    if prec == "f64":
        assert rel.max() < 1e-12
    else:
        assert rel[1:].max() < 1e-6      # FP32 storage of q, volti and rho_old


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_adjoint_parity_and_dot_product(grid_factory, prec):
    grid = grid_factory(12, 20)
    w, st, tr = tracer_case(grid, ntr=2)
    prm = abi.make_params(prec, num_scalar=2, dtsm=600.0)
    g, o = _gpu(grid, prm), Oracle(grid, prm)
    rng = np.random.default_rng(5)
    dq = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    dq[0] = 0
    lam[0] = 0
    lam[:, 0] = 0
    res = {}
    for name, be in (("gpu", g), ("oracle", o)):
        be.upload_state(st)
        be.upload_scalar(1, dq)
        be.upload_scalar(2, tr[1])
        be.scalar_transport_rk(w["rho_old"])
        Ldq = be.download_scalar(1)
        be.upload_scalar_adj(1, lam)
        be.upload_scalar_adj(2, lam)
        be.scalar_transport_rk_adj()
        res[name] = (Ldq, be.download_scalar_adj(1))
    e = rel_linf(res["gpu"][1], res["oracle"][1])
    print(f"{prec} adjoint rel Linf vs oracle {e:.3e}")
    assert e < (1e-11 if prec == "f64" else 5e-6)
    Ldq, LTlam = res["gpu"]
    a = float(np.dot(Ldq[1:, 1:].ravel().astype(np.float64), lam[1:, 1:].ravel()))
    b = float(np.dot(dq[1:].ravel().astype(np.float64), LTlam[1:].ravel()))
    print(f"{prec} dot-product test {a:.10e} {b:.10e} rel {abs(a - b) / abs(a):.3e}")
    assert abs(a - b) / abs(a) < 1e-6


def test_tracer_implicit_column_parity(grid_factory):
    """Large vertical mass fluxes force ff_implic > 0 (the CFL > 1 implicit branch)."""
    grid = grid_factory(10, 20)
    w, st, tr = tracer_case(grid, dt=600.0, ntr=2, wnoise=40.0)
    prm = abi.make_params("f64", num_scalar=2, dtsm=600.0)
    g, o = _gpu(grid, prm), Oracle(grid, prm)
    for be in (g, o):
        load_tracer_case(be, w, st, tr)
        be.scalar_transport_rk(w["rho_old"])
    rep_g, rep_o = g.cfl_report(), o.cfl_report()
    print("cfl gpu", rep_g, "oracle", rep_o)
    assert rep_o[3] > 0, "case did not trigger the implicit branch"
    assert rep_g[3] == rep_o[3]
    assert abs(rep_g[0] - rep_o[0]) / rep_o[0] < 1e-6
    for n in (1, 2):
        e = rel_linf(g.download_scalar(n), o.download_scalar(n))
        print("implicit tracer", n, e)
        assert e < 1e-10
    # adjoint through the implicit branch: dot-product test
    rng = np.random.default_rng(9)
    dq = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    dq[0] = 0
    lam[0] = 0
    lam[:, 0] = 0
    g.upload_scalar(1, dq)
    g.scalar_transport_rk(w["rho_old"])
    Ldq = g.download_scalar(1)
    g.upload_scalar_adj(1, lam)
    g.scalar_transport_rk_adj()
    LTlam = g.download_scalar_adj(1)
    a = float(np.dot(Ldq[1:, 1:].ravel().astype(np.float64), lam[1:, 1:].ravel()))
    b = float(np.dot(dq[1:].ravel().astype(np.float64), LTlam[1:].ravel()))
    print("implicit dot", a, b)
    assert abs(a - b) / abs(a) < 1e-6


@pytest.mark.parametrize("prec", ["f32", "mixed", "f64"])
def test_dynamics_step_parity(grid_factory, prec):
    grid = grid_factory(12, 20)
    jet = cases.jet_state(grid)
    prm = abi.make_params(prec, num_scalar=1, dtsm=600.0)
    g, o = _gpu(grid, prm), Oracle(grid, prm)
    for be in (g, o):
        be.upload_state(jet)
        be.tend0()
        be.prog_wrtv_rk()
        be.prog_wrtv_rk()
    fg, fo = dyn_fields(g, grid), dyn_fields(o, grid)
    tol_big = 1e-11 if prec == "f64" else 1e-5
    tol_small = 1e-9 if prec == "f64" else 1e-4
    for f in DYN_FIELDS:
        e = rel_linf(fg[f], fo[f])
        print(f"{prec} {f:6s} rel Linf {e:.3e}  rel L2 {rel_l2(fg[f], fo[f]):.3e}")
        tol = tol_small if f in ("wmc", "wc", "wmasc") else tol_big
        if prec == "f64" and f == "wc":
            tol = 1e-7       # wmf/rho near the model top, where wmf is small
        assert e < tol, (f, e)


def test_dynamics_with_tracers_mass(grid_factory):
    """prog_wrtv_rk then scalar_transport_rk with the accumulated vmasc/wmasc: dry mass
    and tracer mass are conserved (FP64 mode)."""
    grid = grid_factory(12, 20)
    jet = cases.jet_state(grid)
    prm = abi.make_params("f64", num_scalar=2, dtsm=300.0)
    g = _gpu(grid, prm)
    g.upload_state(jet)
    g.upload_scalar(1, np.full((grid.mwa, grid.mza), 1.0, np.float32))
    g.upload_scalar(2, cases.cosine_bell(grid))
    m0 = g.mass_sums()
    for _ in range(3):
        g.tend0()
        g.prog_wrtv_rk()
        g.timeavg_momsc()
        g.scalar_transport_rk()
    m1 = g.mass_sums()
    rel = np.abs((m1 - m0) / m0)
    print("dyn+tracer mass rel", rel)
    assert rel.max() < 1e-12
    q1 = g.download_scalar(1)[1:, 1:]
    assert np.abs(q1 - 1.0).max() < 1e-12


def test_relabelled_mesh_gives_identical_results(grid_factory):
    """Renumbering is exact: relabelling the input W/V/M points (same neighbour order and
    coefficient bits) yields bit-identical tracer and dynamics results."""
    from helpers import PermutedGrid
    base = grid_factory(10, 16)
    pg = PermutedGrid(base, seed=4)
    w, st, tr = tracer_case(base, ntr=2)
    jet = cases.jet_state(base)
    outs = []
    for grid, perm in ((base, None), (pg, pg)):
        fw = (lambda a, kind="w": a) if perm is None else perm.fwd
        bk = (lambda a, kind="w": a) if perm is None else perm.back
        prm = abi.make_params("f32", num_scalar=2, dtsm=600.0)
        g = _gpu(grid, prm)
        s2 = abi.FortranState(base)
        for name, kind, _ in abi.STATE_FIELDS:
            s2.arrays[name][...] = fw(jet.arrays[name], kind)
        s2.arrays["vmasc"][...] = fw(w["vmasc"], "v")
        s2.arrays["wmasc"][...] = fw(w["wmasc"], "w")
        g.upload_state(s2)
        for n, q in enumerate(tr, start=1):
            g.upload_scalar(n, fw(q))
        g.scalar_transport_rk(fw(w["rho_old"]))
        g.prog_wrtv_rk()
        res = abi.FortranState(base)
        g.download_state(res)
        outs.append([bk(g.download_scalar(2))] + [bk(res.arrays[f], k) for f, k in
                                                  (("thil", "w"), ("vmc", "v"), ("rho", "w"), ("wmc", "w"))])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_invalid_config_is_rejected(grid_factory):
    from olam_amd.gpu import Context, OlamGpuError
    grid = grid_factory(6, 10)
    prm = abi.make_params("f32", num_scalar=1)
    prm.nrk_scal = 4
    with pytest.raises(OlamGpuError):
        Context(grid, prm)
