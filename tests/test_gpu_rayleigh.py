"""GPU parity of the acoustic step with Rayleigh friction near the model top
(mem_rayf.f90:41-220, prog_wrtv_rk.f90:46-52,219-303, vort_div_damp.f90:37-185)
against the oracle, through the C-ABI: every kind on together with each divergence
damping level (the Rayleigh divergence profile shares its buffers with the higher-order
damping), the namelist of the reference's test build (build_olam_test/OLAMIN:444-458:
W, divergence and mixing friction with sixth-order damping), and the call-order and
input checks of the boundary.  Tolerances as in test_gpu_parity.py, except the FP64
vertical fluxes (measured rounding sensitivity, below).
"""

import numpy as np
import pytest

from olam_amd import abi, cases, mesh as M
from olam_amd.gpu import Context
from helpers import DYN_FIELDS, dyn_fields, rel_linf
from oracle import Oracle

pytestmark = pytest.mark.gpu

ALL = {"rayf": (20000.0, 0.3, 1.0), "rayfw": (21000.0, 0.5, 1.0),
       "rayfdiv": (19000.0, 0.15, 1.0), "rayfmix": (22000.0, 0.25, 2.0)}
OLAMIN = {"rayf": (20000.0, 0.0, 1.0), "rayfw": (20000.0, 0.5, 1.0),
          "rayfdiv": (20000.0, 0.15, 1.0), "rayfmix": (20000.0, 0.25, 1.0)}


def _grid(nlev=20):
    g = abi.FortranGrid(M.build_mesh(12), M.vertical_grid(nlev))
    st = cases.jet_state(g)
    g.set_sounding(*cases.reference_sounding(g, st))
    rng = np.random.default_rng(6)
    st.arrays["thil"][1:] += rng.standard_normal(st.arrays["thil"][1:].shape).astype(np.float32)
    st.arrays["wmc"][1:, 1:-1] = 0.05 * rng.standard_normal(st.arrays["wmc"][1:, 1:-1].shape)
    return g, st


@pytest.mark.parametrize("prec,divl,rf,nlev", [("f64", 3, "all", 20), ("f32", 3, "all", 20), ("f64", 2, "all", 20),
                                               ("f64", 1, "all", 20), ("f64", 0, "all", 20),
                                               ("f64", 3, "olamin", 20), ("f32", 3, "olamin", 20),
                                               ("f64", 3, "all", 70), ("f32", 3, "olamin", 70)])
def test_rayleigh_dynamics_parity(prec, divl, rf, nlev):
    grid, st = _grid(nlev)
    prm = abi.make_params(prec, num_scalar=1, dtsm=60.0, divh_damp_level=divl,
                          divh_damp_fact=0.12 if rf == "olamin" else 0.1,
                          rayf=ALL if rf == "all" else OLAMIN)
    print(f"rayleigh {rf} {prec} divh level {divl} L={nlev}", flush=True)
    out = {}
    for name, be in (("gpu", Context(grid, prm, device=0)), ("oracle", Oracle(grid, prm))):
        be.upload_state(st)
        be.init_wrtv_rk()
        be.tend0()
        for _ in range(3):
            be.prog_wrtv_rk()
        out[name] = dyn_fields(be, grid)
        be.close()
    tol_big = 1e-11 if prec == "f64" else 1e-5
    # the W column solve (FP64 PCR on the GPU, Thomas in the oracle) amplifies rounding
    # in the small vertical fluxes: in the oracle alone, scaling rho by 1 + 1e-15 moves
    # wmc, wc and wmasc of these cases by up to 6.7e-9 after the three steps
    tol_small = 2e-8 if prec == "f64" else 1e-4
    for f in DYN_FIELDS:
        e = rel_linf(out["gpu"][f], out["oracle"][f])
        print(f"  {f:6s} rel Linf {e:.3e}")
        tol = tol_small if f in ("wmc", "wc", "wmasc") else tol_big
        if prec == "f64" and f == "wc":
            tol = 1e-7
        if prec == "f64" and nlev > 63 and f in ("wmc", "wmasc", "vze"):
            tol = 1e-7     # tall columns: the 128-row PCR (test_gpu_tall_columns.py)
        assert e < tol, (f, e)


def test_rayleigh_changes_the_step():
    """The friction terms reach the GPU result: the same step with them off differs."""
    grid, st = _grid()
    res = {}
    for name, rf in (("on", ALL), ("off", None)):
        prm = abi.make_params("f64", num_scalar=1, dtsm=60.0, rayf=rf)
        g = Context(grid, prm, device=0)
        g.upload_state(st)
        g.init_wrtv_rk()
        g.tend0()
        g.prog_wrtv_rk()
        res[name] = dyn_fields(g, grid)
        g.close()
    for f in ("thil", "wmc", "vmc", "vxe"):
        assert not np.array_equal(res["on"][f], res["off"][f]), f


def test_rayleigh_call_order_and_inputs():
    grid, st = _grid()
    prm = abi.make_params("f32", num_scalar=1, dtsm=60.0, rayf=ALL)
    g = Context(grid, prm, device=0)
    g.upload_state(st)
    with pytest.raises(RuntimeError, match="init_wrtv_rk"):
        g.prog_wrtv_rk()
    g.close()
    # thil/vmc friction needs the reference sounding
    bare = abi.FortranGrid(grid.mesh, grid.vg, grid.cm)
    with pytest.raises(RuntimeError):
        Context(bare, prm, device=0)
    # without initial == 1 thil/vmc friction is off and init_wrtv_rk is not required
    prm0 = abi.make_params("f32", num_scalar=1, dtsm=60.0, rayf=ALL, initial=0)
    g = Context(bare, prm0, device=0)
    g.upload_state(st)
    g.prog_wrtv_rk()
    g.close()
