"""ctypes binding of libolamgpu.so (the C-ABI in include/olamgpu.h).

This is the host-side mirror of the reference interface for Python callers: the
method names follow the Fortran procedures they replace (``prog_wrtv_rk``,
``scalar_transport_rk``, ``tend0``, ``timeavg_momsc``, ``mass_sums``).  There is no
CPU fallback: if the HIP library is missing or no GPU is visible, construction
raises.

A process that also uses torch must import torch before the library is loaded (the
first Context or lib() call): torch ships its own HIP runtime, which its libraries
request as libamdhip64.so while this library requests the soname libamdhip64.so.7.
Loaded first, torch's runtime carries that soname and this library binds to it; loaded
after, torch starts a second HIP runtime in the process, which finds no GPU.
olam_amd.adjoint imports torch first.  The library does not import torch itself: with
OMP_PROC_BIND set (bench.py sets it for the CPU baseline), importing torch binds the
calling thread to one core, and the library's threaded host copies would share it.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .build import LIB

FIELD = {"scalar": 1, "scalar_adj": 2, "rho": 3, "rho_old": 4, "vmasc": 5, "wmasc": 6, "thil": 7,
         "vmc": 8, "wmc": 9, "press": 10}

_lib = None


def lib():
    """Load libolamgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: run __graft_entry__.build() (the HIP path has no fallback)")
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.olamgpu_create.argtypes = [vp, vp, C.c_int32, C.POINTER(vp)]
        L.olamgpu_destroy.argtypes = [vp]
        L.olamgpu_last_error.argtypes = [vp]
        L.olamgpu_last_error.restype = C.c_char_p
        for name in ("olamgpu_init_wrtv_rk", "olamgpu_tend0", "olamgpu_prog_wrtv_rk", "olamgpu_timeavg_momsc",
                     "olamgpu_scalar_transport_rk_adj"):
            getattr(L, name).argtypes = [vp]
        L.olamgpu_upload_state.argtypes = [vp, vp]
        L.olamgpu_comp_alpha_press.argtypes = [vp, vp, vp, vp]
        L.olamgpu_set_turbulence.argtypes = [vp, vp, vp, vp, vp, vp]
        L.olamgpu_set_nudging.argtypes = [vp, vp, vp, vp]
        L.olamgpu_download_state.argtypes = [vp, vp]
        L.olamgpu_upload_scalar.argtypes = [vp, C.c_int32, vp, vp]
        L.olamgpu_download_scalar.argtypes = [vp, C.c_int32, vp]
        L.olamgpu_upload_scalar_adj.argtypes = [vp, C.c_int32, vp]
        L.olamgpu_download_scalar_adj.argtypes = [vp, C.c_int32, vp]
        L.olamgpu_scalar_transport_rk.argtypes = [vp, vp]
        L.olamgpu_mass_sums.argtypes = [vp, vp]
        L.olamgpu_cfl_report.argtypes = [vp, vp]
        L.olamgpu_copy_field.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_int32, C.POINTER(C.c_int32)]
        L.olamgpu_last_call_ms.argtypes = [vp, C.POINTER(C.c_double)]
        L.olamgpu_set_scalar_pdef.argtypes = [vp, C.c_int32, C.c_int32]
        L.olamgpu_tracer_maxmin.argtypes = [vp, vp]
        L.olamgpu_mass_sums_noalloc.argtypes = [vp, vp]
        L.olamgpu_set_scalar_sgsmix.argtypes = [vp, C.c_int32, C.c_int32]
        L.olamgpu_checkpoint_bytes.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
        L.olamgpu_checkpoint_save.argtypes = [vp, C.c_int32, vp]
        L.olamgpu_checkpoint_restore.argtypes = [vp, C.c_int32, vp]
        L.olamgpu_set_profiling.argtypes = [vp, C.c_int32]
        L.olamgpu_kernel_profile.argtypes = [vp, C.c_char_p, C.c_int32]
        _lib = L
    return _lib


EXPORTS = ["olamgpu_create", "olamgpu_destroy", "olamgpu_last_error", "olamgpu_upload_state",
           "olamgpu_download_state", "olamgpu_upload_scalar", "olamgpu_download_scalar", "olamgpu_init_wrtv_rk",
           "olamgpu_tend0", "olamgpu_prog_wrtv_rk", "olamgpu_timeavg_momsc", "olamgpu_scalar_transport_rk",
           "olamgpu_scalar_transport_rk_adj", "olamgpu_upload_scalar_adj", "olamgpu_download_scalar_adj",
           "olamgpu_mass_sums", "olamgpu_cfl_report", "olamgpu_copy_field", "olamgpu_last_call_ms",
           "olamgpu_set_profiling", "olamgpu_kernel_profile", "olamgpu_checkpoint_bytes",
           "olamgpu_checkpoint_save", "olamgpu_checkpoint_restore", "olamgpu_set_scalar_pdef",
           "olamgpu_comp_alpha_press", "olamgpu_set_turbulence", "olamgpu_set_nudging",
           "olamgpu_set_scalar_sgsmix", "olamgpu_tracer_maxmin", "olamgpu_mass_sums_noalloc"]

CKPT_STATE, CKPT_TAPE = 1, 2


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OlamGpuError(RuntimeError):
    pass


class Context:
    """One olamgpu_ctx: a grid, its device state, and the hot-path calls."""

    def __init__(self, grid: abi.FortranGrid, params: abi.Params, device: int = -1):
        self.grid = grid
        self.params = params
        self.h = C.c_void_p()
        rc = lib().olamgpu_create(C.byref(grid.desc), C.byref(params), int(device), C.byref(self.h))
        if rc:
            raise OlamGpuError(f"olamgpu_create failed ({rc}): {lib().olamgpu_last_error(None).decode()}")

    def _chk(self, rc, what):
        if rc:
            msg = lib().olamgpu_last_error(self.h).decode()
            raise OlamGpuError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.h:
            lib().olamgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state -------------------------------------------------------------------
    def upload_state(self, st, only=None):
        d = st.desc(only)
        self._chk(lib().olamgpu_upload_state(self.h, C.byref(d)), "upload_state")

    def download_state(self, st, only=None):
        d = st.desc(only)
        self._chk(lib().olamgpu_download_state(self.h, C.byref(d)), "download_state")

    def _wfield(self, a, dt, name):
        """A host W field in the Fortran layout, (mwa, mza) contiguous, or ValueError
        (the library reads exactly that many elements)."""
        a = np.ascontiguousarray(a, dtype=dt)
        if a.shape != (self.grid.mwa, self.grid.mza):
            raise ValueError(f"{name} must be (mwa, mza) = {(self.grid.mwa, self.grid.mza)}, got {a.shape}")
        return a

    def upload_scalar(self, n, q, t=None):
        """Tracer n (var_p) and, if given, its long-step tendency (var_t); either may be
        None (a NULL pointer: the device copy of that array is kept)."""
        q = None if q is None else self._wfield(q, np.float32, "q")
        t = None if t is None else self._wfield(t, np.float32, "t")
        self._chk(lib().olamgpu_upload_scalar(self.h, n, _p(q), _p(t)), "upload_scalar")

    def download_scalar(self, n):
        out = np.zeros((self.grid.mwa, self.grid.mza), np.float32)
        self._chk(lib().olamgpu_download_scalar(self.h, n, _p(out)), "download_scalar")
        return out

    def set_scalar_pdef(self, n, pdef):
        self._chk(lib().olamgpu_set_scalar_pdef(self.h, int(n), int(bool(pdef))), "set_scalar_pdef")

    def set_scalar_sgsmix(self, n, on):
        self._chk(lib().olamgpu_set_scalar_sgsmix(self.h, int(n), int(bool(on))), "set_scalar_sgsmix")

    def upload_scalar_adj(self, n, lam):
        lam = self._wfield(lam, np.float32, "lambda")
        self._chk(lib().olamgpu_upload_scalar_adj(self.h, n, _p(lam)), "upload_scalar_adj")

    def download_scalar_adj(self, n):
        out = np.zeros((self.grid.mwa, self.grid.mza), np.float32)
        self._chk(lib().olamgpu_download_scalar_adj(self.h, n, _p(out)), "download_scalar_adj")
        return out

    # hot path ------------------------------------------------------------------
    def init_wrtv_rk(self):
        self._chk(lib().olamgpu_init_wrtv_rk(self.h), "init_wrtv_rk")

    def tend0(self):
        self._chk(lib().olamgpu_tend0(self.h), "tend0")

    def prog_wrtv_rk(self):
        self._chk(lib().olamgpu_prog_wrtv_rk(self.h), "prog_wrtv_rk")

    def timeavg_momsc(self):
        self._chk(lib().olamgpu_timeavg_momsc(self.h), "timeavg_momsc")

    def set_nudging(self, wmanud=None, vmanud=None, rhot_nud=None):
        """mem_nudge fields: wmanud (mwa, mza), vmanud (mva, mza) FP32, rhot_nud (mwa, mza)
        FP64; None leaves a field as it was."""
        g = self.grid
        a = [None if x is None else np.ascontiguousarray(x, dtype=dt)
             for x, dt in ((wmanud, np.float32), (vmanud, np.float32), (rhot_nud, np.float64))]
        for x, n, name in zip(a, (g.mwa, g.mva, g.mwa), ("wmanud", "vmanud", "rhot_nud")):
            if x is not None and x.shape != (n, g.mza):
                raise ValueError(f"{name} must be ({n}, {g.mza}), got {x.shape}")
        self._chk(lib().olamgpu_set_nudging(self.h, *(_p(x) for x in a)), "set_nudging")

    def set_turbulence(self, vkh, vkm, khtop, kmtop, akm_sfc):
        """mem_turb fields of the PBL scheme: vkh, vkm (mwa, mza) FP32 at W levels, khtop,
        kmtop (mwa) int32, akm_sfc (mwa) FP32; then comp_horiz_k on the device."""
        g = self.grid
        f = [np.ascontiguousarray(x, dtype=np.float32) for x in (vkh, vkm, akm_sfc)]
        i = [np.ascontiguousarray(x, dtype=np.int32) for x in (khtop, kmtop)]
        if f[0].shape != (g.mwa, g.mza) or f[1].shape != (g.mwa, g.mza):
            raise ValueError("vkh and vkm must be (mwa, mza)")
        if f[2].shape != (g.mwa,) or i[0].shape != (g.mwa,) or i[1].shape != (g.mwa,):
            raise ValueError("khtop, kmtop and akm_sfc must be (mwa,)")
        self._chk(lib().olamgpu_set_turbulence(self.h, _p(f[0]), _p(f[1]), _p(i[0]), _p(i[1]), _p(f[2])),
                  "set_turbulence")

    def comp_alpha_press(self, rr_v, rr_w, theta):
        """Moist alpha_press, pwfac, pvfac from (mwa, mza) FP32 rr_v, rr_w, theta."""
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (rr_v, rr_w, theta)]
        for x in a:
            if x.shape != (self.grid.mwa, self.grid.mza):
                raise ValueError("rr_v, rr_w and theta must be (mwa, mza)")
        self._chk(lib().olamgpu_comp_alpha_press(self.h, *(_p(x) for x in a)), "comp_alpha_press")

    def scalar_transport_rk(self, rho_old=None):
        r = None if rho_old is None else self._wfield(rho_old, np.float32, "rho_old")
        self._chk(lib().olamgpu_scalar_transport_rk(self.h, _p(r)), "scalar_transport_rk")

    def scalar_transport_rk_adj(self):
        self._chk(lib().olamgpu_scalar_transport_rk_adj(self.h), "scalar_transport_rk_adj")

    # diagnostics -----------------------------------------------------------------
    def mass_sums(self):
        out = np.zeros(1 + self.params.num_scalar)
        self._chk(lib().olamgpu_mass_sums(self.h, _p(out)), "mass_sums")
        return out

    def tracer_maxmin(self):
        """(num_scalar, 2) float32: max and min of each tracer over the prognostic levels
        (tracer_maxmin, check_nans.F90:164-203)."""
        out = np.zeros(2 * max(self.params.num_scalar, 1), np.float32)
        self._chk(lib().olamgpu_tracer_maxmin(self.h, _p(out)), "tracer_maxmin")
        return out[: 2 * self.params.num_scalar].reshape(-1, 2)

    def cfl_report(self):
        out = np.zeros(4)
        self._chk(lib().olamgpu_cfl_report(self.h, _p(out)), "cfl_report")
        return out

    def last_call_ms(self):
        ms = C.c_double()
        lib().olamgpu_last_call_ms(self.h, C.byref(ms))
        return ms.value

    def set_profiling(self, on=True):
        lib().olamgpu_set_profiling(self.h, int(bool(on)))

    def kernel_profile(self):
        """{kernel: (launches, total_ms)} accumulated since set_profiling(True)."""
        buf = C.create_string_buffer(1 << 16)
        lib().olamgpu_kernel_profile(self.h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    # checkpoints (device buffers owned by the caller) ------------------------------
    def checkpoint_bytes(self, kind):
        n = C.c_int64()
        self._chk(lib().olamgpu_checkpoint_bytes(self.h, int(kind), C.byref(n)), "checkpoint_bytes")
        return n.value

    def checkpoint_save(self, kind, dev_ptr):
        self._chk(lib().olamgpu_checkpoint_save(self.h, int(kind), C.c_void_p(dev_ptr)), "checkpoint_save")

    def checkpoint_restore(self, kind, dev_ptr):
        self._chk(lib().olamgpu_checkpoint_restore(self.h, int(kind), C.c_void_p(dev_ptr)), "checkpoint_restore")

    # device-resident access --------------------------------------------------------
    def copy_field(self, name, index, dev_ptr, direction):
        eb = C.c_int32()
        self._chk(lib().olamgpu_copy_field(self.h, FIELD[name], int(index), C.c_void_p(dev_ptr), int(direction),
                                           C.byref(eb)), "copy_field")
        return eb.value
