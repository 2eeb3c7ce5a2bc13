"""Multi-step adjoint windows on the GPU: checkpoint/restore, the torch.autograd
harness, central finite differences, the dot-product test, and a reference gradient
from the CPU oracle (its own per-step adjoint, replayed from the initial state for each
step, with no checkpointing).

The window's tracer map L (dynamics advancing, tracers passive) is linear for fixed
initial dynamics, so J = 1/2 sum w (q(T) - q*)^2 is quadratic in q(0) and the central
difference (J(q + e d) - J(q - e d)) / 2e equals grad J . d up to rounding alone.
"""

import numpy as np
import pytest
import torch

from olam_amd import abi, cases, mesh as M
from olam_amd.gpu import CKPT_STATE, CKPT_TAPE, Context
from olam_amd.adjoint import TransportWindow, transport
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _setup(nxp, nlev, ntr, prec, dt, seed=11):
    grid = abi.FortranGrid(M.build_mesh(nxp), M.vertical_grid(nlev))
    st = cases.jet_state(grid)
    prm = abi.make_params(prec, num_scalar=ntr, dtsm=dt)
    rng = np.random.default_rng(seed)
    q0 = np.stack([1.0 + 0.5 * rng.standard_normal((grid.mwa - 1, grid.mza)) for _ in range(ntr)])
    qs = np.stack([1.0 + 0.5 * rng.standard_normal((grid.mwa - 1, grid.mza)) for _ in range(ntr)])
    return grid, st, prm, q0, qs


def _weights(ctx, grid, dtype):
    """rho * volt of the final state on prognostic levels (k >= 2), zero on level 1,
    normalised to sum 1."""
    rdt = torch.float32 if ctx.params.precision == 0 else torch.float64
    rho = torch.empty((grid.mwa - 1, grid.mza), dtype=rdt, device="cuda")
    ctx.copy_field("rho", 0, rho.data_ptr(), 0)
    w = rho.double() * torch.from_numpy(grid.arrays["volt"][1:]).cuda()
    w[:, 0] = 0.0
    return w / w.sum()


def _ctx(grid, st, prm):
    ctx = Context(grid, prm, device=0)
    ctx.init_wrtv_rk()
    ctx.upload_state(st)
    return ctx


def test_checkpoint_restore_is_bit_exact():
    grid, st, prm, q0, _ = _setup(10, 16, 2, "f32", 120.0)
    ctx = _ctx(grid, st, prm)
    for n in range(2):
        ctx.upload_scalar(n + 1, grid.to_f(q0[n].astype(np.float32)))

    def step():
        ctx.tend0(); ctx.prog_wrtv_rk(); ctx.timeavg_momsc(); ctx.scalar_transport_rk()

    step()
    buf = torch.empty(ctx.checkpoint_bytes(CKPT_STATE), dtype=torch.uint8, device="cuda")
    ctx.checkpoint_save(CKPT_STATE, buf.data_ptr())
    qs = [ctx.download_scalar(n + 1) for n in range(2)]
    step(); step()
    a = abi.FortranState(grid)
    ctx.download_state(a)
    qa = [ctx.download_scalar(n + 1) for n in range(2)]
    ctx.checkpoint_restore(CKPT_STATE, buf.data_ptr())
    for n in range(2):
        ctx.upload_scalar(n + 1, qs[n])
    step(); step()
    b = abi.FortranState(grid)
    ctx.download_state(b)
    for name in ("thil", "rho", "press", "wmc", "wc", "vmc", "vc", "vxe", "vmasc", "wmasc"):
        assert np.array_equal(a.arrays[name], b.arrays[name]), name
    for n in range(2):
        assert np.array_equal(qa[n], ctx.download_scalar(n + 1))
    ctx.close()


def _fd_check(grid, st, prm, q0, qs, nsteps, every, ndir, eps, seed=3):
    ctx = _ctx(grid, st, prm)
    win = TransportWindow(ctx, nsteps, every=every)
    dt = win.dtype
    q0t = torch.tensor(q0, dtype=dt, device="cuda", requires_grad=True)
    qst = torch.tensor(qs, dtype=torch.float64, device="cuda")
    qT = transport(q0t, win)
    w = _weights(ctx, grid, dt)
    J = 0.5 * (w * (qT.double() - qst) ** 2).sum()
    J.backward()
    g = q0t.grad.double()
    assert win.recomputed_steps == nsteps            # dynamics re-run once in total

    def Jof(q):
        with torch.no_grad():
            return float(0.5 * (w * (win.forward(q).double() - qst) ** 2).sum())

    gen = torch.Generator(device="cpu").manual_seed(seed)
    errs = []
    for _ in range(ndir):
        d = torch.randn(q0.shape, generator=gen, dtype=torch.float64).cuda()
        d[:, :, 0] = 0.0
        fd = (Jof((q0t.detach().double() + eps * d).to(dt)) - Jof((q0t.detach().double() - eps * d).to(dt))) / (2 * eps)
        gd = float((g * d).sum())
        errs.append(abs(fd - gd) / abs(gd))
    ctx.close()
    return np.array(errs), win


def test_window_gradient_matches_central_differences_f64():
    grid, st, prm, q0, qs = _setup(12, 16, 2, "f64", 120.0)
    errs, win = _fd_check(grid, st, prm, q0, qs, nsteps=6, every=2, ndir=8, eps=1e-3)
    print(f"f64 FD rel errors max {errs.max():.3e} median {np.median(errs):.3e}")
    assert errs.max() < 1e-8


def test_window_c4_gradient_fd_64_directions():
    """NXP=160, L=50, 4 FP32 tracers, 20 steps, checkpoint every 5, 64 seeded Gaussian
    directions, eps = 1e-3.  In FP32 the central difference carries the rounding noise
    of two 20-step forward runs divided by 2 eps; the bound below is that noise level."""
    grid, st, prm, q0, qs = _setup(160, 50, 4, "f32", 60.0)
    errs, win = _fd_check(grid, st, prm, q0, qs, nsteps=20, every=5, ndir=64, eps=1e-3)
    print(f"C4 f32 FD rel errors max {errs.max():.3e} median {np.median(errs):.3e}; "
          f"checkpoint+tape memory {win.device_bytes() / 2**30:.2f} GiB")
    assert np.median(errs) < 1e-3 and errs.max() < 1e-2


def test_window_dot_product():
    grid, st, prm, q0, _ = _setup(16, 20, 3, "f32", 120.0)
    ctx = _ctx(grid, st, prm)
    win = TransportWindow(ctx, 8, every=3)
    rng = np.random.default_rng(9)
    dq = torch.tensor(rng.standard_normal(q0.shape), dtype=torch.float32, device="cuda")
    lam = torch.tensor(rng.standard_normal(q0.shape), dtype=torch.float32, device="cuda")
    lam[:, :, 0] = 0.0
    Ldq = win.forward(dq)
    LTlam = win.backward(lam)
    a = float((Ldq.double() * lam.double()).sum())
    b = float((dq.double() * LTlam.double()).sum())
    print(f"window dot-product {a:.10e} {b:.10e} rel {abs(a - b) / abs(a):.3e}")
    assert abs(a - b) / abs(a) < 1e-6
    ctx.close()


def test_window_gradient_matches_oracle_reference():
    """The oracle has no checkpoints: for each step i (last to first) it replays steps
    0..i from the initial state so its tape is step i's, then runs one adjoint step."""
    grid, st, prm, q0, qs = _setup(8, 12, 2, "f64", 120.0)
    nsteps = 4
    ctx = _ctx(grid, st, prm)
    win = TransportWindow(ctx, nsteps, every=2)
    lam = np.random.default_rng(4).standard_normal(q0.shape)
    lam[:, :, 0] = 0.0
    win.forward(torch.tensor(q0, dtype=torch.float64, device="cuda"))
    g_gpu = win.backward(torch.tensor(lam, dtype=torch.float64, device="cuda")).cpu().numpy()
    ctx.close()

    lam_f = [grid.to_f(lam[n].astype(np.float32)) for n in range(2)]
    cur = lam_f
    for i in reversed(range(nsteps)):
        o = Oracle(grid, prm)
        o.upload_state(st)
        for n in range(2):
            o.upload_scalar(n + 1, grid.to_f(q0[n].astype(np.float32)))
        for _ in range(i + 1):
            o.tend0(); o.prog_wrtv_rk(); o.timeavg_momsc(); o.scalar_transport_rk()
        for n in range(2):
            o.upload_scalar_adj(n + 1, cur[n])
        o.scalar_transport_rk_adj()
        cur = [o.download_scalar_adj(n + 1) for n in range(2)]
        o.close()
    # the oracle's adjoint buffers are FP32 host arrays: compare at FP32 resolution
    for n in range(2):
        ref = cur[n][1:].astype(np.float64)
        err = np.abs(g_gpu[n] - ref).max() / np.abs(ref).max()
        print(f"tracer {n}: GPU window gradient vs oracle reference rel Linf {err:.3e}")
        assert err < 1e-6
