"""Build helpers: the HIP library (gfx950), the oracle (test infrastructure) and the
Fortran ISO_C_BINDING interface module.

Everything is built in-tree so the shared objects travel with the repository:
  olam_amd/lib/libolamgpu.so      product: HIP kernels + C-ABI (include/olamgpu.h)
  olam_amd/lib/olamgpu_iface.mod  Fortran interface module (+ .o) when amdflang exists
  olam_amd/lib/libolamio.so       HDF5 gridfile / history storage (C libhdf5 of the image)
  oracle/build/liboracle.so       CPU restatement used only by tests and the CPU baseline
"""

from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libolamgpu.so")
FORTRAN = os.path.join(PKG, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 (v_pk_*_f32) is kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build_hip(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(CSRC, "olamgpu.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


COPY = os.path.join(LIBDIR, "hbm_copy")


def build_copy(verbose=False):
    """The plain-copy program used for bandwidth and counter calibration."""
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", COPY, os.path.join(CSRC, "hbm_copy.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return COPY


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return os.path.join(ROOT, "oracle", "build", "liboracle.so")


def flang():
    for cand in ("/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/flang", shutil.which("amdflang") or ""):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran():
    """Compile the ISO_C_BINDING interface module and the Fortran test driver."""
    fc = flang()
    if fc is None:
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    iface = os.path.join(FORTRAN, "olamgpu_iface.f90")
    driver = os.path.join(FORTRAN, "olamgpu_driver.f90")
    subprocess.run([fc, "-c", "-O2", "-J", LIBDIR, "-o", os.path.join(LIBDIR, "olamgpu_iface.o"), iface],
                   check=True)
    exe = os.path.join(LIBDIR, "olamgpu_fdriver")
    if os.path.exists(driver):
        subprocess.run([fc, "-O2", "-I", LIBDIR, "-o", exe, driver, os.path.join(LIBDIR, "olamgpu_iface.o"),
                        "-L", LIBDIR, "-lolamgpu", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN"],
                       check=True)
    build_shim_driver(fc)
    return exe


SHIM_DRIVER = os.path.join(LIBDIR, "olam_shim_driver")


def build_shim_driver(fc):
    """The drop-in shim (olam_amd/fortran/olam_shim.f90: modules wrtv_rk and
    scalar_transport) compiled against the test stand-ins of OLAM's host modules
    (tests/fortran/host_modules.f90) with the test host tests/fortran/shim_driver.f90.
    Modules and objects go to olam_amd/lib/shim/, apart from the product's module."""
    tdir = os.path.join(ROOT, "tests", "fortran")
    srcs = [os.path.join(tdir, "host_modules.f90"), os.path.join(FORTRAN, "olam_shim.f90"),
            os.path.join(tdir, "shim_driver.f90")]
    if not all(os.path.exists(x) for x in srcs):
        return None
    out = os.path.join(LIBDIR, "shim")
    os.makedirs(out, exist_ok=True)
    objs = []
    for src in srcs:
        obj = os.path.join(out, os.path.basename(src)[:-4] + ".o")
        subprocess.run([fc, "-c", "-O2", "-J", out, "-I", LIBDIR, "-o", obj, src], check=True)
        objs.append(obj)
    subprocess.run([fc, "-O2", "-o", SHIM_DRIVER, *objs, os.path.join(LIBDIR, "olamgpu_iface.o"),
                    "-L", LIBDIR, "-lolamgpu", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN"], check=True)
    return SHIM_DRIVER


def build_io(verbose=False):
    """The HDF5 storage layer of the gridfile / history I/O (olam_amd/gridfile.py)."""
    from . import gridfile
    if not os.path.exists(os.path.join(gridfile.HDF5_PREFIX, "include", "hdf5.h")):
        return None
    return gridfile.build(verbose)


def build_all(verbose=False):
    build_hip(verbose)
    build_copy(verbose)
    build_oracle()
    build_io(verbose)
    if os.path.exists(os.path.join(FORTRAN, "olamgpu_iface.f90")):
        build_fortran()
