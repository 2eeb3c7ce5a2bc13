"""FP32-bulk and reference-precision drift against FP64, and mass conservation, over
many long steps.

    python -m olam_amd.drift [--nxp 320] [--nlev 60] [--ntr 10] [--steps 100] [--dt 30]

Runs the same initial state (the balanced zonal jet with a 1 K theta bump and
log-normal tracers) through N long steps of tend0, prog_wrtv_rk, timeavg_momsc and
scalar_transport_rk on the GPU in the FP64 mode (which matches the CPU oracle to 1e-12
on the parity cases), the FP32-bulk mode and the reference-precision mode (FP64 rho and
press), then reports per prognostic field and per tracer the relative L2 and Linf
differences of each reduced mode from FP64, and the relative change of the dry-air and
tracer masses (sum rho volt, sum rho volt q) of each run.  Prints one JSON object.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi, cases, mesh as M
from .gpu import Context

FIELDS = ("thil", "rho", "press", "wmc", "wc", "vmc", "vc", "vxe", "vye", "vze")


def run(grid, st0, tracers, prec, nsteps, dt):
    ctx = Context(grid, abi.make_params(prec, num_scalar=len(tracers), dtsm=dt), device=0)
    ctx.init_wrtv_rk()
    ctx.upload_state(st0)
    for n, q in enumerate(tracers):
        ctx.upload_scalar(n + 1, q)
    m0 = ctx.mass_sums()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.tend0()
        ctx.prog_wrtv_rk()
        ctx.timeavg_momsc()
        ctx.scalar_transport_rk()
    sec = time.perf_counter() - t0
    m1 = ctx.mass_sums()
    out = abi.FortranState(grid)
    ctx.download_state(out)
    qs = [ctx.download_scalar(n + 1) for n in range(len(tracers))]
    ctx.close()
    return out, qs, m0, m1, sec


def rel(a, b):
    a = np.asarray(a, np.float64)[1:, 1:]
    b = np.asarray(b, np.float64)[1:, 1:]
    d = a - b
    nb2 = np.linalg.norm(b)
    nbi = np.abs(b).max()
    return (float(np.linalg.norm(d) / nb2) if nb2 > 0 else float(np.linalg.norm(d)),
            float(np.abs(d).max() / nbi) if nbi > 0 else float(np.abs(d).max()))


def drift(nxp=320, nlev=60, ntr=10, nsteps=100, dt=30.0, modes=("f32", "mixed"), grid=None, log=None):
    if grid is None:
        grid = abi.FortranGrid(M.build_mesh(nxp), M.vertical_grid(nlev))
    st0 = cases.jet_state(grid)
    tr = cases.lognormal_tracers(grid, ntr)
    runs = {}
    for prec in ("f64",) + tuple(modes):
        if log:
            log(f"[drift] {prec}: {nsteps} steps")
        runs[prec] = run(grid, st0, tr, prec, nsteps, dt)
    hi, qhi, m0hi, m1hi, shi = runs["f64"]
    res = {"workload": f"NXP={nxp} ({grid.mwa - 1} columns), L={grid.mza - 1}, "
                       f"{ntr} tracers, {nsteps} steps of {dt} s",
           "mass_change": {"f64": ((m1hi - m0hi) / m0hi).tolist()}, "seconds": {"f64": shi}}
    for prec in modes:
        lo, qlo, m0lo, m1lo, slo = runs[prec]
        r = {"fields": {}, "tracers": {}}
        for f in FIELDS:
            l2, li = rel(lo.arrays[f], hi.arrays[f])
            r["fields"][f] = {"rel_l2": l2, "rel_linf": li}
        for n in range(ntr):
            l2, li = rel(qlo[n], qhi[n])
            r["tracers"][f"q{n + 1}"] = {"rel_l2": l2, "rel_linf": li}
        r["max_tracer_rel_l2"] = max(v["rel_l2"] for v in r["tracers"].values())
        res[prec] = r
        res["mass_change"][prec] = ((m1lo - m0lo) / m0lo).tolist()
        res["seconds"][prec] = slo
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nxp", type=int, default=320)
    ap.add_argument("--nlev", type=int, default=60)
    ap.add_argument("--ntr", type=int, default=10)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--dt", type=float, default=30.0)
    a = ap.parse_args()
    print(json.dumps(drift(a.nxp, a.nlev, a.ntr, a.steps, a.dt), indent=1), flush=True)


if __name__ == "__main__":
    main()
