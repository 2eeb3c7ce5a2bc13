"""Benchmark of the RK3 dynamical core on one MI355X.

One step = one long step of the hot path with nacoust = 1 (so dtsm = dtlm):
  tend0 -> prog_wrtv_rk (acoustic RK3: W/V momentum, rho, rho*thil, press, damping,
  vel_t3d) -> timeavg_momsc -> scalar_transport_rk (RK3, all tracers) ->
  scalar_transport_rk_adj (adjoint of that tracer step, all tracers).
Workload: global icosahedral-hexagonal mesh NXP=320 (1,024,002 columns), 60 levels,
10 tracers with log-normal mixing ratios, dt = 30 s, synthetic jet state (no
datasets).  Metric: cell-updates/s = columns x levels / seconds per step.

Beside the metric the line carries the roofline of the dominant kernel, the step's
fraction of the HBM roofline, host-resident (PCIe-inclusive) rates, the reference-
precision mode, one GPU's share of the 70-level, 20-tracer configuration (NXP=320:
a quarter of NXP=640's columns, two waves per column) and the CPU baseline.

Prints one JSON line.  --gpus N > 1 exits with a message: this build targets one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads: one per core of this process's share, bound close
# (read by the OpenMP runtime when it starts, so set before anything loads it)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

PEAK_HBM = 8.0e12          # MI355X HBM3E, bytes/s


def alg_bytes(ntr, prec):
    """Algorithmic bytes per cell-level per RK3 step (dynamics, tracer fwd, tracer adj).

    Dynamics 772 B (FP32 bulk) or 856 B (FP64 rho/press); tracers 120 + 48 n each for
    the forward and the adjoint (per-stage shared 24 B + 16 B per tracer, 3 stages,
    plus 48 B of setup); all-FP64 doubles everything."""
    dyn = 856 if prec == "mixed" else 772
    tr = 120 + 48 * ntr
    if prec == "f64":
        return 2 * dyn, 2 * tr, 2 * tr
    return dyn, tr, tr


def stage_alg_bytes(name, ntr, prec):
    """Algorithmic bytes per cell-level of the RK stage that one launch of the named
    kernel carries out, from the per-stage byte model: a tracer stage (all tracers,
    forward or adjoint) reads rhos, volti, wmasc and vmasc on 3 faces (24 B) and per
    tracer reads scp, scp0, sct and writes scp (16 B), so 24 + 16 n (the long step's
    48 B of setup is not a stage's); the dynamics' W-column phase 132 (152 with FP64
    rho/press), its V phase 92 (100) and its velocity phase 28.  A stage that runs as
    two kernels credits each with the whole stage's bytes.  This is what the headline
    roofline uses, whatever the kernel moves besides."""
    tr = 24.0 + 16.0 * ntr
    mixed = prec == "mixed"
    per = {
        "k_tr_grad": tr, "k_tr_update": tr, "k_tr_stage": tr,
        "k_tr_adj0": tr, "k_tr_adj1": tr, "k_tr_adj01": tr, "k_tr_adj2": tr,
        "k_dyn_w": 152 if mixed else 132,
        "k_dyn_v": 100 if mixed else 92,
        "k_dyn_vel": 28,
    }.get(name)
    if per is None:
        return None
    return 2 * per if prec == "f64" else per


def kernel_alg_bytes(name, ntr, prec):
    """Bytes per cell-level of every array one launch of the named kernel must read or
    write, counted once (V-face fields as 3 faces per cell), including intermediates
    that the stage model does not count (gradients, partial adjoints)."""
    w = 8 if prec == "f64" else 4
    per = {
        "k_tr_grad": 3 * ntr,                 # q -> gx, gy
        "k_tr_update": 6 + 5 * ntr,           # rho, volti, wmasc, vmasc; q, q0, gx, gy -> q
        "k_tr_adj0": 6 + 5 * ntr,             # vmasc, wmasc, volti, rho; mu, lam0 -> lam0, nu, lpart
        "k_tr_adj1": 3 + 5 * ntr,             # vmasc; nu, lpart -> lpart, lgx, lgy
        "k_tr_adj01": 6 + 6 * ntr,            # vmasc, wmasc, volti, rho; mu, lam0 -> lam0, lpart, lgx, lgy
        "k_tr_adj2": 4 * ntr,                 # lpart, lgx, lgy -> lq
        "k_tr_stage": 6 + 3 * ntr,            # rho, volti, wmasc, vmasc; q, q0 -> q
        # thil, vxe, vye, vze, their 8 gradients, 4 short-step tendencies, rho, volt (FP64),
        # wmc, volti, alpha_press, rho0, rth0, pwfac, volwi, wmc0, wc, arw; vmc, arv on 3
        # faces per cell -> press, rho, thil, vm[xyz]et_rk, wmc, wc
        "k_dyn_w": 43,
        "k_dyn_v": 23,                        # P2
        "k_dyn_vel": 7,                       # P3
    }.get(name)
    if per is None:
        return None
    extra = {"k_dyn_w": 20, "k_dyn_v": 8}.get(name, 0) if prec == "mixed" else 0
    if prec == "f64" and name == "k_dyn_w":
        extra = -8                           # volt is one FP64 word, counted as two FP32 words
    return per * w + extra


def committed_traffic(kernel, nxp, nlev, ntr, prec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of
    this workload (FETCH_SIZE and WRITE_SIZE runs, copy-calibrated), if there is one."""
    path = os.path.join(ROOT, "profiles", f"hbm_counters_nxp{nxp}_L{nlev}_tr{ntr}_{prec}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:
            try:
                read_gb, write_gb = float(cells[2]), float(cells[4])
            except ValueError:
                continue
            return (read_gb + write_gb) * 1e9, os.path.relpath(path, ROOT)
    return None, None


def host_cpu():
    """CPU model and the number of cores this process may run on (its share of the host)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(grid, ntr, prec, dt, threads, max_steps=5, time_box=60.0, warmup=2):
    """Oracle (CPU restatement, OpenMP) on the metric workload: `warmup` untimed steps,
    then timed steps until max_steps or the time box (at least 2); the median step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from olam_amd import abi, cases
    jet = cases.jet_state(grid)
    tr = cases.lognormal_tracers(grid, ntr)
    used = O.set_threads(threads)
    prm = abi.make_params(prec, num_scalar=ntr, dtsm=dt)
    o = O.Oracle(grid, prm)
    o.upload_state(jet)
    for n in range(ntr):
        o.upload_scalar(n + 1, tr[n])
        o.upload_scalar_adj(n + 1, tr[n])

    def step():
        o.tend0()
        o.prog_wrtv_rk()
        o.timeavg_momsc()
        o.scalar_transport_rk()
        o.scalar_transport_rk_adj()

    for _ in range(warmup):                  # the first allocates the scratch
        step()
    times = []
    t_start = time.perf_counter()
    while len(times) < max_steps and (len(times) < 2 or time.perf_counter() - t_start < time_box):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    o.close()
    sec = float(sorted(times)[len(times) // 2])
    nw, nlev = grid.mwa - 1, grid.mza - 1
    model, _ = host_cpu()
    return {"value": nw * nlev / sec, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "cpu_model": model, "seconds_per_step": sec, "steps_timed": len(times),
            "sample": f"oracle (C++ restatement, OpenMP {os.environ.get('OMP_PROC_BIND')}/"
                      f"{os.environ.get('OMP_PLACES')}, -O3 -march={O.ISA} -ffp-contract=off) on {nw} columns, "
                      f"L={nlev}, {ntr} tracers, precision {prec}: {warmup} warm-up steps, then the median of "
                      f"{len(times)} timed steps of dynamics + tracer forward + tracer adjoint "
                      f"(time box {time_box:g} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nxp", type=int, default=320)
    ap.add_argument("--nlev", type=int, default=60)
    ap.add_argument("--ntr", type=int, default=10)
    ap.add_argument("--dt", type=float, default=30.0)
    ap.add_argument("--prec", default="f32", choices=["f32", "mixed", "f64"])
    ap.add_argument("--cpu-nxp", type=int, default=None, help="CPU baseline mesh (default: the workload's)")
    ap.add_argument("--cpu-prec", default="mixed", help="CPU baseline precision (default: the reference's)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core this process may run on")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-time-box", type=float, default=60.0)
    ap.add_argument("--cpu-1thread-nxp", type=int, default=40,
                    help="mesh of the additional 1-thread CPU figure (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c5-steps", type=int, default=5,
                    help="timed steps of one GPU's share of the 70-level, 20-tracer configuration (0: skip)")
    ap.add_argument("--c5-nxp", type=int, default=320)
    ap.add_argument("--mixed-steps", type=int, default=5,
                    help="also time this many steps in the reference-precision mode (0: skip)")
    ap.add_argument("--no-host-resident", action="store_true",
                    help="skip timing long steps whose state and tracers start and end in host memory")
    ap.add_argument("--no-copy", action="store_true", help="skip the achievable-copy-bandwidth probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU decomposition is described, not built)"}))
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    from olam_amd import abi, cases, mesh
    from olam_amd.build import COPY
    from olam_amd.gpu import Context

    copy_bw = None
    if not args.no_copy and os.path.exists(COPY):
        # achievable HBM copy bandwidth, measured by a child process before this one
        # touches the GPU
        import subprocess
        r = subprocess.run([COPY, str(1 << 30), "10"], capture_output=True, text=True, timeout=120)
        if r.returncode == 0:
            copy_bw = json.loads(r.stdout.strip().splitlines()[-1])["copy_GBps"]
            log(f"[bench] achievable copy bandwidth {copy_bw:.0f} GB/s")

    t = time.perf_counter()
    m = mesh.build_mesh(args.nxp)
    vg = mesh.vertical_grid(args.nlev)
    grid = abi.FortranGrid(m, vg)
    log(f"[bench] mesh NXP={args.nxp}: {m.nw} columns, {m.nv} faces, {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    jet = cases.jet_state(grid)
    tr = cases.lognormal_tracers(grid, args.ntr)
    log(f"[bench] synthetic state + {args.ntr} tracers: {time.perf_counter() - t:.1f}s")
    cells = m.nw * args.nlev

    def make_ctx(prec):
        c = Context(grid, abi.make_params(prec, num_scalar=args.ntr, dtsm=args.dt), device=0)
        c.upload_state(jet)
        for n in range(args.ntr):
            c.upload_scalar(n + 1, tr[n])
            c.upload_scalar_adj(n + 1, tr[n])
        return c

    def stepper(c):
        def step():
            c.tend0()
            c.prog_wrtv_rk()
            c.timeavg_momsc()
            c.scalar_transport_rk()
            c.scalar_transport_rk_adj()
        return step

    t = time.perf_counter()
    ctx = make_ctx(args.prec)
    step = stepper(ctx)
    log(f"[bench] device setup (renumbering, checks, uploads): {time.perf_counter() - t:.1f}s")
    for _ in range(args.warmup):
        step()
    log("[bench] warm-up done")
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()          # every C-ABI call returns after synchronising its stream
    sec = (time.perf_counter() - t0) / args.steps
    prof = ctx.kernel_profile()
    ctx.set_profiling(False)
    value = cells / sec
    log(f"[bench] {sec * 1e3:.2f} ms/step, {value:.3e} cell-updates/s")
    kern_ms = {k: v[1] / args.steps for k, v in prof.items()}
    for k, (n, ms) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        log(f"[bench]   {k:16s} {n // args.steps:3d} launches/step  {ms / args.steps:8.3f} ms/step  "
            f"{ms / n:8.3f} ms/launch")

    host = None
    if not args.no_host_resident:
        # the drop-in as the Fortran shim drives it when the host owns the state.  "shim":
        # per long step the long-step tendencies (thilt, vmxet, vmyet, vmzet, vmt) and the
        # tracer tendencies go up, the fields the host's physics reads (thil, rho, press,
        # vxe, vye, vze, wc) and the tracers come down.  "full": the whole state and the
        # tracers go up and come down around every step.
        st_out = abi.FortranState(grid)
        zero_t = np.zeros((grid.mwa, grid.mza), np.float32)
        tend = ("thilt", "vmxet", "vmyet", "vmzet", "vmt")
        phys = ("thil", "rho", "press", "vxe", "vye", "vze", "wc")
        st_tend = abi.FortranState(grid)

        def shim_step():
            ctx.tend0()
            ctx.upload_state(st_tend, only=tend)
            ctx.prog_wrtv_rk()
            ctx.timeavg_momsc()
            for n in range(args.ntr):
                ctx.upload_scalar(n + 1, None, zero_t)
            ctx.scalar_transport_rk()
            ctx.scalar_transport_rk_adj()
            ctx.download_state(st_out, only=phys)
            for n in range(args.ntr):
                ctx.download_scalar(n + 1)

        def full_step():
            ctx.upload_state(jet)
            for n in range(args.ntr):
                ctx.upload_scalar(n + 1, tr[n])
            step()
            ctx.download_state(st_out)
            for n in range(args.ntr):
                ctx.download_scalar(n + 1)

        def timed(fn):
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return sorted(ts)[1]

        w4, v4, w8 = grid.mwa * grid.mza * 4, grid.mva * grid.mza * 4, grid.mwa * grid.mza * 8
        shim_up = 4 * w4 + v4 + args.ntr * w4
        shim_down = 5 * w4 + 2 * w8 + args.ntr * w4           # rho and press are real(r8)
        nv_f = sum(1 for _, k, _ in abi.STATE_FIELDS if k == "v")
        nw8 = sum(1 for _, k, d in abi.STATE_FIELDS if k == "w" and d == np.float64)
        nw4 = len(abi.STATE_FIELDS) - nv_f - nw8
        full_b = nw4 * w4 + nw8 * w8 + nv_f * v4 + args.ntr * w4
        ssec = timed(shim_step)
        fsec = timed(full_step)
        host = {"shim": {"ms_per_step": ssec * 1e3, "transfer_ms": (ssec - sec) * 1e3,
                         "bytes_up": shim_up, "bytes_down": shim_down,
                         "GBps": (shim_up + shim_down) / max(ssec - sec, 1e-9) / 1e9,
                         "value": cells / ssec},
                "full": {"ms_per_step": fsec * 1e3, "transfer_ms": (fsec - sec) * 1e3,
                         "bytes_up": full_b, "bytes_down": full_b,
                         "GBps": 2 * full_b / max(fsec - sec, 1e-9) / 1e9,
                         "value": cells / fsec},
                "what": "host-resident long steps (median of 3): shim = the Fortran shim's per-step "
                        "transfers, full = the whole state and tracers up and down every step; "
                        "bytes are the host arrays' (Fortran layout)"}
        log(f"[bench] host-resident shim step {ssec * 1e3:.1f} ms ({(ssec - sec) * 1e3:.1f} ms transfers, "
            f"{host['shim']['GBps']:.1f} GB/s); full {fsec * 1e3:.1f} ms ({host['full']['GBps']:.1f} GB/s)")
    ctx.close()

    dom = max(prof.items(), key=lambda kv: kv[1][1])[0]
    n_launch, tot_ms = prof[dom]
    per_launch_s = tot_ms / n_launch * 1e-3
    sb = stage_alg_bytes(dom, args.ntr, args.prec)
    kb = kernel_alg_bytes(dom, args.ntr, args.prec)
    roof = None
    if sb is not None:
        ach = sb * cells / per_launch_s
        traffic, tsrc = committed_traffic(dom, args.nxp, args.nlev, args.ntr, args.prec)
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "traffic": traffic, "traffic_source": tsrc,
                "frac_of_copy": (ach / (copy_bw * 1e9)) if copy_bw else None,
                "alg_bytes_per_launch": sb * cells, "alg_bytes_basis": "per-stage bytes (tracer stage 24 + 16 n; dynamics phases 132, 92, 28)",
                "avg_launch_ms": per_launch_s * 1e3,
                "frac_kernel_arrays": (kb * cells / per_launch_s / PEAK_HBM) if kb else None}
    bd, bf, ba = alg_bytes(args.ntr, args.prec)
    step_frac = (bd + bf + ba) * cells / sec / PEAK_HBM

    mixed = None
    if args.mixed_steps > 0 and args.prec == "f32":
        c2 = make_ctx("mixed")
        s2 = stepper(c2)
        s2()
        t0 = time.perf_counter()
        for _ in range(args.mixed_steps):
            s2()
        msec = (time.perf_counter() - t0) / args.mixed_steps
        c2.close()
        mixed = {"ms_per_step": msec * 1e3, "value": cells / msec, "steps": args.mixed_steps,
                 "dtype": "fp32 + fp64 rho/press (the reference's precision map)"}
        log(f"[bench] reference-precision mode: {msec * 1e3:.2f} ms/step")

    c5 = None
    if args.c5_steps > 0:
        # one GPU's share of NXP=640 x 70 levels x 20 tracers (a quarter of its columns),
        # at its real column shape: 70 levels run as two waves per column
        t = time.perf_counter()
        g5 = abi.FortranGrid(mesh.build_mesh(args.c5_nxp), mesh.vertical_grid(70))
        jet5 = cases.jet_state(g5)
        tr5 = cases.lognormal_tracers(g5, 20)
        c = Context(g5, abi.make_params("f32", num_scalar=20, dtsm=args.dt), device=0)
        c.upload_state(jet5)
        for n in range(20):
            c.upload_scalar(n + 1, tr5[n])
            c.upload_scalar_adj(n + 1, tr5[n])
        del tr5
        log(f"[bench] 70-level, 20-tracer share NXP={args.c5_nxp}: set-up {time.perf_counter() - t:.1f}s")
        s5 = stepper(c)
        for _ in range(2):
            s5()
        c.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(args.c5_steps):
            s5()
        sec5 = (time.perf_counter() - t0) / args.c5_steps
        p5 = c.kernel_profile()
        c.close()
        cells5 = (g5.mwa - 1) * 70
        c5 = {"workload": f"prog_wrtv_rk + scalar_transport_rk + adjoint, NXP={args.c5_nxp} ({g5.mwa - 1} columns, "
                          f"one GPU's share of NXP=640), L=70 (two waves per column), 20 tracers, FP32",
              "ms_per_step": sec5 * 1e3, "value": cells5 / sec5, "unit": "cell-updates/s", "steps": args.c5_steps,
              "kernel_ms_per_step": {k: round(v[1] / args.c5_steps, 3)
                                     for k, v in sorted(p5.items(), key=lambda kv: -kv[1][1])[:6]}}
        log(f"[bench] 70-level, 20-tracer share: {sec5 * 1e3:.2f} ms/step, {cells5 / sec5:.3e} cell-updates/s")

    cpu = cpu1 = None
    if not args.no_cpu:
        threads = args.cpu_threads or host_cpu()[1]
        cgrid = grid if args.cpu_nxp in (None, args.nxp) else \
            abi.FortranGrid(mesh.build_mesh(args.cpu_nxp), mesh.vertical_grid(args.nlev))
        log(f"[bench] CPU baseline (oracle) on {threads} threads ...")
        cpu = cpu_baseline(cgrid, args.ntr, args.cpu_prec, args.dt, threads, args.cpu_steps, args.cpu_time_box)
        log(f"[bench] CPU {cpu['value']:.3e} cell-updates/s on {cpu['cores']} threads "
            f"({cpu['seconds_per_step']:.2f} s/step)")
        if args.cpu_1thread_nxp:
            g1 = abi.FortranGrid(mesh.build_mesh(args.cpu_1thread_nxp), mesh.vertical_grid(args.nlev))
            cpu1 = cpu_baseline(g1, args.ntr, args.cpu_prec, args.dt, 1, 5, 20.0)
            log(f"[bench] CPU 1 thread {cpu1['value']:.3e} cell-updates/s")
    out = {
        "metric": "cell-updates/s per RK3 step (fwd+adjoint) at NXP=320, L=60; % HBM roofline",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32": "fp32 (fp64 column solve and mass sums)", "mixed": "fp32 + fp64 rho/press",
                  "f64": "fp64"}[args.prec],
        "data": "synthetic (seeded jet state, log-normal tracers)",
        "config": {"workload": f"prog_wrtv_rk + scalar_transport_rk + adjoint, NXP={args.nxp} "
                               f"({m.nw} columns), L={args.nlev}, {args.ntr} tracers, dt={args.dt:g}s, RK3",
                   "columns": m.nw, "levels": args.nlev, "tracers": args.ntr, "precision": args.prec,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_cell_level": bd + bf + ba, "frac": step_frac,
                          "achieved_GBps": (bd + bf + ba) * cells / sec / 1e9,
                          "copy_GBps": copy_bw,
                          "frac_of_copy": ((bd + bf + ba) * cells / sec / (copy_bw * 1e9)) if copy_bw else None},
        "kernel_ms_per_step": {k: round(v, 4) for k, v in kern_ms.items()},
        "reference_precision": mixed,
        "l70_20tracer_share": c5,
        "host_resident": host,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
