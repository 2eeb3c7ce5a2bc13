"""Oracle pins of two terms of scalar_transport_rk that act when the PBL scheme or the
nudging is on (CPU only):

* Horizontal tracer diffusion into the tendency (scalar_transport_rk.F90:244-274,
  do_sgsmix tracers, akhodx/khtopv of comp_horiz_k): the face fluxes are
  antisymmetric, so a tracer's mass is unchanged by them (to rounding) while its
  variance falls; a constant tracer gets nothing; do_sgsmix = .false. switches it off;
  the adjoint (which carries the transpose of the diffusion) passes the dot-product
  test, and fails it by orders of magnitude without the transpose.
* The nudged final-stage density rhof = rho - dtl rhot_nud (:120-133) with
  nud_preserve_mix_ratio: when the dynamics added dtl rhot_nud to the density, the
  tracer mass measured with rhof is conserved by the flux-form step to rounding; with
  nud_preserve_mix_ratio off it is not.
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from helpers import tracer_case
from oracle import Oracle
from test_turbulence import turbulence

DT = 600.0


@pytest.fixture(scope="module")
def grid():
    return abi.FortranGrid(M.build_mesh(8), M.vertical_grid(14))


def strong_turbulence(grid, st, seed):
    """The seeded PBL fields with diffusivities large enough that comp_horiz_k's
    stability bound (pbl_driver.f90:215-230) sets akhodx on most faces."""
    vkh, vkm, khtop, kmtop, akm = turbulence(grid, st, seed=seed, drag=False)
    return (vkh * np.float32(1e6)).astype(np.float32), vkm, khtop, kmtop, akm


def _setup(grid, noise=1.0, ntr=2, sgs=(1, 1), turb=True, iscal_monot=0, seed=3):
    w, st, tr = tracer_case(grid, dt=DT, ntr=ntr, noise=noise)
    st.arrays["rho"][:] = w["rho"]
    prm = abi.make_params("f64", num_scalar=ntr, dtsm=DT, iscal_monot=iscal_monot)
    o = Oracle(grid, prm)
    o.upload_state(st)
    if turb:
        o.set_turbulence(*strong_turbulence(grid, st, seed))
    for n, on in enumerate(sgs, start=1):
        o.set_scalar_sgsmix(n, on)
    return o, w, st, tr


def _mass(grid, rho, q):
    v = grid.arrays["volt"][1:, 1:]
    return float((rho[1:, 1:] * v * q[1:, 1:].astype(np.float64)).sum())


def test_diffusion_conserves_mass_and_reduces_variance(grid):
    """No wind, constant density: the step is the diffusion alone."""
    rng = np.random.default_rng(7)
    st = abi.FortranState(grid)
    st.arrays["rho"][:] = 1.0
    q = (1.0 + 0.3 * rng.standard_normal((grid.mwa, grid.mza))).astype(np.float32)
    out = {}
    for on in (0, 1):
        o = Oracle(grid, abi.make_params("f64", num_scalar=2, dtsm=DT))
        o.upload_state(st)
        o.set_turbulence(*strong_turbulence(grid, st, 2))
        o.set_scalar_sgsmix(1, on)
        o.set_scalar_sgsmix(2, on)
        o.upload_scalar(1, q)
        o.upload_scalar(2, np.ones_like(q))
        o.scalar_transport_rk(np.ones((grid.mwa, grid.mza), np.float32))
        out[on] = (o.download_scalar(1), o.download_scalar(2))
        o.close()
    rho = np.ones((grid.mwa, grid.mza))
    q1 = out[1][0]
    assert np.array_equal(out[0][0][1:, 1:], q[1:, 1:])     # without diffusion nothing moves
    assert not np.array_equal(q1[1:, 1:], q[1:, 1:])
    assert abs(_mass(grid, rho, q1) / _mass(grid, rho, q) - 1) < 1e-7   # FP32 tracer I/O
    v = grid.arrays["volt"][1:, 1:]
    var = lambda a: float((v * (a[1:, 1:].astype(np.float64) - 1.0) ** 2).sum())   # noqa: E731
    print(f"variance {var(q1) / var(q):.3f} of the initial")
    assert var(q1) < 0.9 * var(q)
    assert np.abs(out[1][1][1:, 1:] - 1.0).max() == 0.0     # constant tracer


def _dot_test(grid, sgs_adj=True):
    o, w, st, tr = _setup(grid, ntr=1, sgs=(1,))
    rng = np.random.default_rng(2)
    dq = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    dq[0] = 0
    lam[0] = 0
    o.upload_scalar(1, dq)
    o.scalar_transport_rk(w["rho_old"])
    Ldq = o.download_scalar(1)
    if not sgs_adj:
        o.set_scalar_sgsmix(1, 0)         # the adjoint then leaves out the transpose
    o.upload_scalar_adj(1, lam)
    o.scalar_transport_rk_adj()
    LT = o.download_scalar_adj(1)
    a = float(np.dot(Ldq[1:].ravel().astype(np.float64), lam[1:].ravel()))
    b = float(np.dot(dq[1:].ravel().astype(np.float64), LT[1:].ravel()))
    scale = float(np.linalg.norm(Ldq[1:].astype(np.float64)) * np.linalg.norm(lam[1:].astype(np.float64)))
    o.close()
    return abs(a - b) / scale


def test_adjoint_dot_product_with_diffusion(grid):
    good = _dot_test(grid)
    bad = _dot_test(grid, sgs_adj=False)
    print(f"dot-product mismatch {good:.2e}; without the transpose {bad:.2e}")
    assert good < 5e-9 and bad > 1e3 * good


def test_adjoint_refuses_diffusion_with_limiter(grid):
    o, w, st, tr = _setup(grid, ntr=1, sgs=(1,), iscal_monot=1)
    o.upload_scalar(1, tr[0])
    o.scalar_transport_rk(w["rho_old"])
    o.upload_scalar_adj(1, tr[0])
    with pytest.raises(RuntimeError):
        o.scalar_transport_rk_adj()
    o.close()


@pytest.mark.parametrize("preserve_mix", [1, 0])
def test_nudged_density_conserves_tracer_mass(grid, preserve_mix):
    w, st, tr = tracer_case(grid, dt=DT, ntr=2)
    rng = np.random.default_rng(9)
    rt = np.zeros((grid.mwa, grid.mza))
    rt[1:, 1:] = 1e-6 * rng.standard_normal((grid.mwa - 1, grid.mza - 1))
    st.arrays["rho"][:] = w["rho"] + DT * rt           # the dynamics added dtl rhot_nud
    prm = abi.make_params("f64", num_scalar=2, dtsm=DT, nudflag=1, nud_preserve_total_mass=0,
                          nud_preserve_mix_ratio=preserve_mix)
    o = Oracle(grid, prm)
    o.upload_state(st)
    o.set_nudging(None, None, rt)
    for n, q in enumerate(tr, start=1):
        o.upload_scalar(n, q)
    o.scalar_transport_rk(w["rho_old"])
    q1 = o.download_scalar(2)
    o.close()
    m0 = _mass(grid, w["rho_old"].astype(np.float64), tr[1])
    m1 = _mass(grid, w["rho"], q1)
    rel = abs(m1 / m0 - 1)
    print(f"preserve_mix {preserve_mix}: tracer mass change {rel:.2e}")
    # measured 1.3e-9 (the FP32 download of the tracer) and 2.0e-6
    if preserve_mix:
        assert rel < 2e-8
    else:
        assert rel > 2e-7
