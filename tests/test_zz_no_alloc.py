"""olamgpu_mass_sums_noalloc, the allocation-free variant of olamgpu_mass_sums (partial
sums in a static device array), through the C-ABI: two calls in a row on one GPU, the
second on a smaller mesh with fewer tracers, each against olamgpu_mass_sums of the same
context (FP64 sums in another order: 1e-13 relative), and the output written into a
device array the caller owns."""

import ctypes as C

import numpy as np
import pytest
import torch

from olam_amd import abi, cases, mesh as M
from olam_amd.gpu import Context, lib
from helpers import tracer_case

pytestmark = pytest.mark.gpu


def _case(nxp, nlev, ntr, prec):
    grid = abi.FortranGrid(M.build_mesh(nxp), M.vertical_grid(nlev))
    w, st, tr = tracer_case(grid, ntr=ntr)
    g = Context(grid, abi.make_params(prec, num_scalar=ntr, dtsm=600.0), device=0)
    g.upload_state(st)
    for n, q in enumerate(tr, start=1):
        g.upload_scalar(n, q)
    g.scalar_transport_rk(w["rho_old"])
    return g


def _noalloc(g, ntr):
    out = torch.full((1 + ntr,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = lib()
    rc = L.olamgpu_mass_sums_noalloc(g.h, C.c_void_p(out.data_ptr()))
    assert rc == 0, L.olamgpu_last_error(g.h)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_mass_sums_noalloc_matches_mass_sums(prec):
    print(f"mass_sums_noalloc {prec}: NXP=12 L=16 4 tracers, then NXP=8 L=10 2 tracers", flush=True)
    for nxp, nlev, ntr in ((12, 16, 4), (8, 10, 2)):
        g = _case(nxp, nlev, ntr, prec)
        ref = g.mass_sums()
        got = _noalloc(g, ntr)
        g.close()
        assert got.shape == ref.shape and np.isfinite(got).all()
        assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref)), (got, ref)
