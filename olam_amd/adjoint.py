"""Multi-step tracer windows with the hand-written adjoint, checkpointing and a
torch.autograd.Function wrapper.

One step of the window is one long step of OLAM's timestep (omodel/timestep.f90:
190-260): tend0, nacoust x prog_wrtv_rk, timeavg_momsc (the "dynamics", which produce
the mass fluxes vmasc/wmasc and densities a tracer step reads), then
scalar_transport_rk.  For fixed dynamics a tracer step is linear in the tracers (the
upwind choices depend on the winds only), so the window's tracer map L is linear and
its adjoint is the product of the per-step adjoints in reverse order, each run with
that step's winds (the "tape": vmasc, wmasc, rho, rho_old).  With a final-stage
limiter (iscal_monot > 0) the step is linear only with the limiter's scales frozen at
the forward trajectory; the tape then also holds those scales, the checkpoints hold
the tracers, and the backward pass re-runs each tracer step to regenerate them.

Memory: the dynamics state is checkpointed every ``every`` steps during the forward
pass (default ceil(sqrt(N))); the backward pass restores each segment's checkpoint,
re-runs the segment's dynamics once while saving one tape per step, then runs the
tracer adjoint of the segment's steps in reverse.  Peak device memory beyond the
context is ceil(N/every) state checkpoints + every tapes; the dynamics are recomputed
once in total.  All buffers are device tensors; nothing crosses PCIe in the window.

Tracer tensors are dense (num_scalar, N_W, mza) in the original (unrenumbered) point
order with the dummy point removed, on the context's device, float32 (float64 in the
F64 precision mode).
"""

from __future__ import annotations

import math

import torch

from .gpu import CKPT_STATE, CKPT_TAPE, Context


class TransportWindow:
    def __init__(self, ctx: Context, nsteps: int, every: int | None = None, dynamics: bool = True,
                 device: torch.device | str = "cuda"):
        if nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        self.ctx = ctx
        self.nsteps = nsteps
        self.every = every or max(1, math.ceil(math.sqrt(nsteps)))
        self.dynamics = dynamics
        self.device = torch.device(device)
        self.ntr = ctx.params.num_scalar
        self.nw = ctx.grid.mwa - 1
        self.mza = ctx.grid.mza
        self.dtype = torch.float64 if ctx.params.precision == 2 else torch.float32
        # with a final-stage limiter the tracer map is linear only with the scales
        # frozen at the forward trajectory: the backward pass re-runs the tracer steps
        self.limited = ctx.params.iscal_monot > 0
        nseg = math.ceil(nsteps / self.every)
        sb = ctx.checkpoint_bytes(CKPT_STATE)
        tb = ctx.checkpoint_bytes(CKPT_TAPE)
        self.state_bytes, self.tape_bytes = sb, tb
        self.ckpts = [torch.empty(sb, dtype=torch.uint8, device=self.device) for _ in range(nseg)]
        self.tapes = [torch.empty(tb, dtype=torch.uint8, device=self.device)
                      for _ in range(min(self.every, nsteps))]
        # the window always starts from the context's state at construction
        torch.cuda.synchronize(self.device)
        ctx.checkpoint_save(CKPT_STATE, self.ckpts[0].data_ptr())
        self.recomputed_steps = 0

    def device_bytes(self):
        return len(self.ckpts) * self.state_bytes + len(self.tapes) * self.tape_bytes

    # tracer transfers (device to device) ----------------------------------------
    def _put(self, name, q):
        q = q.detach().to(self.device, self.dtype).contiguous()
        if q.shape != (self.ntr, self.nw, self.mza):
            raise ValueError(f"expected shape {(self.ntr, self.nw, self.mza)}, got {tuple(q.shape)}")
        torch.cuda.synchronize(self.device)
        for n in range(self.ntr):
            self.ctx.copy_field(name, n + 1, q[n].data_ptr(), 1)

    def _get(self, name):
        out = torch.empty((self.ntr, self.nw, self.mza), dtype=self.dtype, device=self.device)
        torch.cuda.synchronize(self.device)
        for n in range(self.ntr):
            self.ctx.copy_field(name, n + 1, out[n].data_ptr(), 0)
        return out

    def _dyn(self):
        if not self.dynamics:
            return
        c = self.ctx
        c.tend0()
        for _ in range(c.params.nacoust):
            c.prog_wrtv_rk()
        c.timeavg_momsc()

    # the window -----------------------------------------------------------------
    def forward(self, q0: torch.Tensor) -> torch.Tensor:
        """q(T) of the window from q(0); saves the segment checkpoints."""
        c = self.ctx
        c.checkpoint_restore(CKPT_STATE, self.ckpts[0].data_ptr())
        self._put("scalar", q0)
        if self.limited:
            # the state checkpoints also hold the tracers (the backward pass re-runs the
            # tracer steps): the first one must hold q(0)
            torch.cuda.synchronize(self.device)
            c.checkpoint_save(CKPT_STATE, self.ckpts[0].data_ptr())
        for i in range(self.nsteps):
            if i > 0 and i % self.every == 0:
                c.checkpoint_save(CKPT_STATE, self.ckpts[i // self.every].data_ptr())
            self._dyn()
            c.scalar_transport_rk()
        return self._get("scalar")

    def backward(self, lam_T: torch.Tensor) -> torch.Tensor:
        """dJ/dq(0) from dJ/dq(T) = lam_T (adjoint of forward); needs a prior forward."""
        c = self.ctx
        self._put("scalar_adj", lam_T)
        nseg = len(self.ckpts)
        for s in reversed(range(nseg)):
            i0, i1 = s * self.every, min(self.nsteps, (s + 1) * self.every)
            c.checkpoint_restore(CKPT_STATE, self.ckpts[s].data_ptr())
            for i in range(i0, i1):
                self._dyn()
                if self.limited:
                    # the limiter scales depend on the tracers: re-run the tracer step
                    # (the checkpoint holds the tracers) to tape them
                    c.scalar_transport_rk()
                c.checkpoint_save(CKPT_TAPE, self.tapes[i - i0].data_ptr())
                self.recomputed_steps += 1
            for i in reversed(range(i0, i1)):
                c.checkpoint_restore(CKPT_TAPE, self.tapes[i - i0].data_ptr())
                c.scalar_transport_rk_adj()
        return self._get("scalar_adj")


class TransportWindowFn(torch.autograd.Function):
    """q(T) = L q(0) through a TransportWindow; backward runs the hand-written adjoint."""

    @staticmethod
    def forward(fctx, q0, window):
        fctx.window = window
        return window.forward(q0)

    @staticmethod
    def backward(fctx, grad_out):
        return fctx.window.backward(grad_out), None


def transport(q0: torch.Tensor, window: TransportWindow) -> torch.Tensor:
    return TransportWindowFn.apply(q0, window)
