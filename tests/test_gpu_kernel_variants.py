"""The tracer stage has five kernel arrangements selected when a context is created
(OLAMGPU_TRACER): the default (a gradient pass and an update pass with the gradient
pairs stored side by side, and the adjoint's column and face parts in one pass),
"twopass" (the same with gx and gy in separate arrays), "fused" (one pass per stage that
forms every gradient where it is used, at orders 1 and 2), "split" (the adjoint's
column and face parts as two passes with nu written between them) and "patch"
(LDS-staged patches of columns).  They evaluate
the same expressions in the same order, so their outputs agree to the last bit or to
FMA-contraction rounding; each is checked here against the default, and the default is
checked against the oracle by the parity tests.  At order 3 "patch" runs the gather
kernels (the LDS-staged kernels carry the order-2 reconstruction only).
"""

import os

import numpy as np
import pytest
import torch

from olam_amd import abi
from olam_amd.gpu import Context
from helpers import tracer_case

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-13, "f32": 2e-6}


def _get(g, name, ntr):
    """The tracers (or their adjoints) at full precision, device to device."""
    dt = torch.float64 if g.params.precision == 2 else torch.float32
    out = torch.empty((ntr, g.grid.mwa - 1, g.grid.mza), dtype=dt, device="cuda")
    torch.cuda.synchronize()
    for n in range(ntr):
        g.copy_field(name, n + 1, out[n].data_ptr(), 0)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _rel(a, b):
    return float(np.abs(a[:, 1:] - b[:, 1:]).max() / np.abs(b[:, 1:]).max())


def _run(grid, prm, variant, w, st, tr, lam, wnoise):
    old = os.environ.get("OLAMGPU_TRACER")
    if variant == "default":
        os.environ.pop("OLAMGPU_TRACER", None)
    else:
        os.environ["OLAMGPU_TRACER"] = variant
    try:
        g = Context(grid, prm, device=0)
    finally:
        if old is None:
            os.environ.pop("OLAMGPU_TRACER", None)
        else:
            os.environ["OLAMGPU_TRACER"] = old
    g.upload_state(st)
    for n, q in enumerate(tr, start=1):
        g.upload_scalar(n, q)
    g.scalar_transport_rk(w["rho_old"])
    q = _get(g, "scalar", len(tr))
    for n in range(1, len(tr) + 1):
        g.upload_scalar_adj(n, lam[n - 1])
    g.scalar_transport_rk_adj()
    a = _get(g, "scalar_adj", len(tr))
    g.close()
    return q, a


@pytest.mark.parametrize("prec,order,wnoise", [("f64", 2, 0.01), ("f32", 2, 0.01), ("f64", 1, 0.01),
                                               ("f64", 2, 40.0), ("f64", 3, 0.01), ("f32", 3, 0.01)])
def test_variants_agree(grid_factory, prec, order, wnoise):
    grid = grid_factory(12, 24)
    ntr = 5
    w, st, tr = tracer_case(grid, ntr=ntr, wnoise=wnoise)
    prm = abi.make_params(prec, num_scalar=ntr, dtsm=600.0, order=order)
    rng = np.random.default_rng(11)
    lam = [rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32) for _ in range(ntr)]
    variants = ("twopass", "fused", "split", "patch")
    res = {v: _run(grid, prm, v, w, st, tr, lam, wnoise) for v in ("default",) + variants}
    q0, a0 = res["default"]
    for v in variants:
        q, a = res[v]
        for n in range(ntr):
            eq, ea = _rel(q[n], q0[n]), _rel(a[n], a0[n])
            same = np.array_equal(q[n], q0[n]) and np.array_equal(a[n], a0[n])
            print(f"{prec} order={order} wnoise={wnoise} {v} tracer {n + 1}: forward {eq:.2e} "
                  f"adjoint {ea:.2e} bit-identical {same}")
            assert eq <= TOL[prec] and ea <= TOL[prec]
