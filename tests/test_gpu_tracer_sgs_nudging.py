"""GPU parity, through the C-ABI, of the two tracer-step terms that the PBL scheme and
the nudging switch on: the horizontal diffusion of do_sgsmix tracers into their
tendencies (scalar_transport_rk.F90:244-274) with its transpose in the adjoint, and
the nudged final-stage density rho - dtl rhot_nud (:120-133).  Tolerances as in
test_gpu_parity.py's tracer cases (F32 1e-5, F64 1e-12 relative Linf); also at L=70
(tall columns).
"""

import numpy as np
import pytest

from olam_amd import abi, mesh as M
from olam_amd.gpu import Context
from helpers import rel_linf, tracer_case
from oracle import Oracle
from test_tracer_sgs_nudging import strong_turbulence

pytestmark = pytest.mark.gpu

DT = 600.0


@pytest.mark.parametrize("prec,nlev", [("f32", 16), ("f64", 16), ("f32", 70), ("f64", 70)])
def test_tracer_diffusion_and_adjoint_parity(prec, nlev):
    grid = abi.FortranGrid(M.build_mesh(12), M.vertical_grid(nlev))
    w, st, tr = tracer_case(grid, dt=DT, ntr=3)
    st.arrays["rho"][:] = w["rho"]
    turb = strong_turbulence(grid, st, 3)
    rng = np.random.default_rng(12)
    lam = [rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32) for _ in range(3)]
    prm = abi.make_params(prec, num_scalar=3, dtsm=DT)
    print(f"tracer diffusion {prec} L={nlev}", flush=True)
    out = {}
    for name, be in (("gpu", Context(grid, prm, device=0)), ("oracle", Oracle(grid, prm))):
        be.upload_state(st)
        be.set_turbulence(*turb)
        be.set_scalar_sgsmix(3, 0)
        for n, q in enumerate(tr, start=1):
            be.upload_scalar(n, q)
        be.scalar_transport_rk(w["rho_old"])
        fwd = [be.download_scalar(n) for n in (1, 2, 3)]
        for n in (1, 2, 3):
            be.upload_scalar_adj(n, lam[n - 1])
        be.scalar_transport_rk_adj()
        out[name] = fwd + [be.download_scalar_adj(n) for n in (1, 2, 3)]
        be.close()
    tol = 1e-12 if prec == "f64" else 1e-5
    for i, (a, b) in enumerate(zip(out["gpu"], out["oracle"])):
        e = rel_linf(a, b)
        print(f"  {'fwd' if i < 3 else 'adj'} tracer {i % 3 + 1} rel Linf {e:.3e}")
        assert e < tol, (i, e)


@pytest.mark.parametrize("prec,nlev", [("f32", 16), ("f64", 16), ("f64", 70)])
def test_nudged_density_parity(prec, nlev):
    grid = abi.FortranGrid(M.build_mesh(12), M.vertical_grid(nlev))
    w, st, tr = tracer_case(grid, dt=DT, ntr=2)
    rng = np.random.default_rng(9)
    rt = np.zeros((grid.mwa, grid.mza))
    rt[1:, 1:] = 1e-6 * rng.standard_normal((grid.mwa - 1, grid.mza - 1))
    st.arrays["rho"][:] = w["rho"] + DT * rt
    lam = rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32)
    prm = abi.make_params(prec, num_scalar=2, dtsm=DT, nudflag=1, nud_preserve_total_mass=0)
    print(f"nudged density {prec}", flush=True)
    out = {}
    for name, be in (("gpu", Context(grid, prm, device=0)), ("oracle", Oracle(grid, prm))):
        be.upload_state(st)
        be.set_nudging(None, None, rt)
        for n, q in enumerate(tr, start=1):
            be.upload_scalar(n, q)
        be.scalar_transport_rk(w["rho_old"])
        fwd = be.download_scalar(2)
        be.upload_scalar_adj(2, lam)
        be.scalar_transport_rk_adj()
        out[name] = (fwd, be.download_scalar_adj(2))
        be.close()
    tol = 1e-12 if prec == "f64" else 1e-5
    for a, b in zip(out["gpu"], out["oracle"]):
        e = rel_linf(a, b)
        print(f"  rel Linf {e:.3e}")
        assert e < tol
