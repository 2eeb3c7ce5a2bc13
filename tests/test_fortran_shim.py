"""The drop-in shim olam_amd/fortran/olam_shim.f90 (modules wrtv_rk and
scalar_transport with OLAM's public procedures init_wrtv_rk, prog_wrtv_rk and
scalar_transport_rk), compiled against stand-ins of OLAM's host modules
(tests/fortran/host_modules.f90: mem_grid, mem_ijtabs with the itab_w/itab_v/itab_m
derived types, mem_adv, mem_basic, mem_tend, misc_coms, consts_coms, oname_coms,
var_tables) and driven by tests/fortran/shim_driver.f90 as olam_run / timestep call it.

CPU: the shim and the driver build (build_fortran) and the shim names, in its host
modules' `use` lists, the variables OLAM declares there.
GPU: two long steps with two acoustic steps each through the shim give the fields it
brings back every long step, the whole state after olamgpu_shim_download_all and the
tracers bit-identical to the same steps made from Python through ctypes, without and
with long-step tendencies of the host physics (seeded thilt, vmxet, vmyet, vmzet, vmt
and tracer tendencies, which the shim uploads every long step).
"""

import os
import re
import subprocess

import numpy as np
import pytest

from olam_amd import abi, build, cases, mesh as M
from helpers import write_fortran_case

SHIM = os.path.join(build.FORTRAN, "olam_shim.f90")


def test_shim_uses_reference_module_names():
    text = open(SHIM).read()
    mods = set(re.findall(r"^\s*module\s+(\w+)\s*$", text, re.M))
    assert {"wrtv_rk", "scalar_transport", "olamgpu_host"} <= mods
    pub = re.findall(r"public\s*::\s*([\w, ]+)", text)
    names = {x.strip() for p in pub for x in p.split(",")}
    assert {"init_wrtv_rk", "prog_wrtv_rk", "scalar_transport_rk"} <= names
    assert re.search(r"subroutine\s+scalar_transport_rk\s*\(\s*rho_old\s*\)", text)
    used = set(re.findall(r"use\s+(\w+)\s*,\s*only", text))
    assert {"mem_grid", "mem_ijtabs", "mem_adv", "mem_basic", "mem_tend", "misc_coms",
            "consts_coms", "oname_coms", "var_tables"} <= used


@pytest.mark.gpu
@pytest.mark.parametrize("prec,physics", [("mixed", False), ("f32", False), ("mixed", True)])
def test_shim_long_steps_match_ctypes(tmp_path, prec, physics):
    from olam_amd.gpu import Context
    exe = build.SHIM_DRIVER
    if not os.path.exists(exe):
        pytest.skip("shim driver not built (no amdflang at build time)")
    grid = abi.FortranGrid(M.build_mesh(10), M.vertical_grid(20))
    st = cases.jet_state(grid)
    tr = cases.lognormal_tracers(grid, 3)
    prm = abi.make_params(prec, num_scalar=3, dtsm=120.0, nacoust=2, dtlm=240.0, adjoint_tape=0)
    nsteps = 2
    tends = None
    if physics:   # small seeded tendencies of the host physics, the same every long step
        rng = np.random.default_rng(21)
        for name, scale in (("thilt", 1e-3), ("vmxet", 1e-3), ("vmyet", 1e-3), ("vmzet", 1e-4), ("vmt", 1e-3)):
            a = st.arrays[name]
            a[...] = (scale * rng.standard_normal(a.shape)).astype(np.float32)
            a[0] = 0
        tends = []
        for n in range(3):
            t = (1e-9 * rng.standard_normal((grid.mwa, grid.mza))).astype(np.float32)
            t[0] = 0
            tends.append(t)
    write_fortran_case(str(tmp_path), grid, prm, st, tr, nsteps)
    if tends:
        with open(tmp_path / "fields.txt", "a") as f:
            for n, t in enumerate(tends):
                t.tofile(str(tmp_path / f"scalart{n + 1}.bin"))
                f.write(f"scalart scalart{n + 1} {t.nbytes}\n")
    print(f"shim driver: nxp=10 L=20 ntr=3 {prec} nsteps=2 nacoust=2 physics={physics}", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    ctx = Context(grid, prm, device=0)
    ctx.upload_state(st)
    ctx.init_wrtv_rk()
    for n in range(3):
        ctx.upload_scalar(n + 1, tr[n])
    tend = ("thilt", "vmxet", "vmyet", "vmzet", "vmt")
    for _ in range(nsteps):
        ctx.tend0()
        ctx.upload_state(st, only=tend)
        for _ in range(prm.nacoust):
            ctx.prog_wrtv_rk()
        ctx.timeavg_momsc()
        for n in range(3):
            ctx.upload_scalar(n + 1, None, tends[n] if tends else np.zeros((grid.mwa, grid.mza), np.float32))
        ctx.scalar_transport_rk()
    out = abi.FortranState(grid)
    ctx.download_state(out)
    qs = [ctx.download_scalar(n + 1) for n in range(3)]
    ctx.close()

    def shim(name, like):
        return np.fromfile(str(tmp_path / f"{name}.bin"), dtype=like.dtype).reshape(like.shape)

    for name in ("thil", "rho", "press", "vxe", "vye", "vze", "wc"):
        a = out.arrays[name]
        assert np.array_equal(shim(f"shim_phys_{name}", a)[1:], a[1:]), name
    for name in ("thil", "wmc", "wc", "vxe", "vye", "vze", "rho", "press", "vmc", "vc", "wmasc", "vmasc"):
        a = out.arrays[name]
        assert np.array_equal(shim(f"shim_{name}", a)[1:], a[1:]), name
    for n in range(3):
        assert np.array_equal(shim(f"shim_scalar{n + 1}", qs[n])[1:], qs[n][1:]), n
