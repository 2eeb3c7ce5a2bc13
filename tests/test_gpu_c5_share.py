"""One GPU's share of the largest configuration, at its real shape: NXP=320 (1,024,002
columns, a quarter of NXP=640's 4,096,002, i.e. one GPU of a four-way split of the
columns), 70 levels (tall columns: two wavefronts per column), 20 FP32 tracers, FP32
dynamics with the FP64 column solve, dt = 30 s, over a 40-step window with
ceil(sqrt(40)) = 7-step checkpoint segments (6 state checkpoints, 7 tapes, all in HBM).

* dot-product test of the window, <L dq, lam> against <dq, L^T lam> for seeded
  Gaussian dq and lam, to 1e-6;
* dry-air and tracer mass over the window's forward pass from the seeded log-normal
  tracers (mu = -20, sigma = 2): FP32 storage of rho and q rounds every stored value by
  up to 6e-8 relative each step, and the sums drift by the sum of those roundings
  (6.9e-10 after 100 steps at 60 levels): bound 3e-9;
* the device memory the context, the checkpoints, the tapes and the test's own
  tensors hold is printed.
"""

import time

import numpy as np
import pytest
import torch

from olam_amd import abi, cases, mesh as M
from olam_amd.adjoint import TransportWindow
from olam_amd.gpu import Context

pytestmark = pytest.mark.gpu

NXP, NLEV, NTR, NSTEPS = 320, 70, 20, 40


def _dot(x, y):
    """<x, y> in FP64, one tracer at a time."""
    return sum(float(torch.dot(x[n].reshape(-1).double(), y[n].reshape(-1).double())) for n in range(x.shape[0]))


def test_c5_share_window_dot_product_and_mass():
    print(f"C5 share: NXP={NXP} L={NLEV} {NTR} f32 tracers, {NSTEPS}-step window", flush=True)
    t = time.perf_counter()
    grid = abi.FortranGrid(M.build_mesh(NXP), M.vertical_grid(NLEV))
    st = cases.jet_state(grid)
    tr = cases.lognormal_tracers(grid, NTR)
    prm = abi.make_params("f32", num_scalar=NTR, dtsm=30.0)
    ctx = Context(grid, prm, device=0)
    ctx.init_wrtv_rk()
    ctx.upload_state(st)
    for n in range(NTR):
        ctx.upload_scalar(n + 1, tr[n])
    m0 = ctx.mass_sums()
    win = TransportWindow(ctx, NSTEPS)
    assert win.every == 7 and len(win.ckpts) == 6
    free, total = torch.cuda.mem_get_info()
    print(f"set-up {time.perf_counter() - t:.1f}s; checkpoints + tapes {win.device_bytes() / 1e9:.1f} GB; "
          f"device memory in use {(total - free) / 1e9:.1f} of {total / 1e9:.1f} GB", flush=True)

    # mass over the forward pass
    q0 = torch.from_numpy(np.stack([q[1:] for q in tr])).cuda()
    del tr
    t = time.perf_counter()
    qT = win.forward(q0)
    torch.cuda.synchronize()
    print(f"forward window {time.perf_counter() - t:.1f}s", flush=True)
    m1 = ctx.mass_sums()
    rel = np.abs((m1 - m0) / m0)
    print(f"dry and tracer mass rel change over the window: dry {rel[0]:.2e}, tracers max {rel[1:].max():.2e}")
    assert bool(torch.isfinite(qT[:, :, 1:]).all())
    assert rel.max() < 3e-9, rel
    del q0, qT

    # dot-product test
    gen = torch.Generator(device="cuda").manual_seed(9)
    shape = (NTR, grid.mwa - 1, grid.mza)
    dq = torch.randn(shape, generator=gen, dtype=torch.float32, device="cuda")
    lam = torch.randn(shape, generator=gen, dtype=torch.float32, device="cuda")
    lam[:, :, 0] = 0.0
    t = time.perf_counter()
    Ldq = win.forward(dq)
    LTlam = win.backward(lam)
    torch.cuda.synchronize()
    print(f"forward + backward window {time.perf_counter() - t:.1f}s; "
          f"peak torch memory {torch.cuda.max_memory_allocated() / 1e9:.1f} GB", flush=True)
    a, b = _dot(Ldq, lam), _dot(dq, LTlam)
    print(f"window dot-product {a:.12e} {b:.12e} rel {abs(a - b) / abs(a):.3e}")
    assert abs(a - b) / abs(a) < 1e-6
    assert win.recomputed_steps == NSTEPS
    ctx.close()
