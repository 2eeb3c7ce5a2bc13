"""ctypes loader of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product path (olam_amd.gpu) never imports it.  The oracle takes the same
olamgpu_mesh / olamgpu_params / olamgpu_state descriptors as the product C-ABI, so a
test can feed identical inputs to both.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_V3 = os.path.join(HERE, "build", "liboracle.so")
LIB_V4 = os.path.join(HERE, "build", "liboracle_avx512.so")


def _has_avx512():
    need = {"avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512cd"}
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return need <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


# the x86-64-v4 build where the CPU has AVX-512 (same bits, wider vectors);
# OLAM_ORACLE_ISA=v3 forces the AVX2 build
ISA = "x86-64-v4" if os.environ.get("OLAM_ORACLE_ISA", "") != "v3" and _has_avx512() else "x86-64-v3"
LIB = LIB_V4 if ISA == "x86-64-v4" else LIB_V3


def build(force=False):
    if force or not (os.path.exists(LIB_V3) and os.path.exists(LIB_V4)):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.oracle_create.argtypes = [vp, vp, C.c_int32, C.POINTER(vp)]
        for name in ("oracle_destroy", "oracle_tend0", "oracle_prog_wrtv_rk", "oracle_timeavg_momsc",
                     "oracle_scalar_transport_rk_adj", "oracle_init_wrtv_rk"):
            getattr(L, name).argtypes = [vp]
        L.oracle_upload_state.argtypes = [vp, vp]
        L.oracle_download_state.argtypes = [vp, vp]
        L.oracle_upload_scalar.argtypes = [vp, C.c_int32, vp, vp]
        L.oracle_download_scalar.argtypes = [vp, C.c_int32, vp]
        L.oracle_upload_scalar_adj.argtypes = [vp, C.c_int32, vp]
        L.oracle_download_scalar_adj.argtypes = [vp, C.c_int32, vp]
        L.oracle_scalar_transport_rk.argtypes = [vp, vp]
        L.oracle_mass_sums.argtypes = [vp, vp]
        L.oracle_cfl_report.argtypes = [vp, vp]
        L.oracle_tracer_maxmin.argtypes = [vp, vp]
        L.oracle_set_threads.argtypes = [C.c_int32]
        L.oracle_tridiffo.argtypes = [C.c_int32, vp, vp, vp, vp, vp]
        L.oracle_set_scalar_pdef.argtypes = [vp, C.c_int32, C.c_int32]
        L.oracle_set_scalar_sgsmix.argtypes = [vp, C.c_int32, C.c_int32]
        L.oracle_set_frozen_limiter.argtypes = [vp, C.c_int32]
        L.oracle_tape_scales.argtypes = [vp, C.c_int32, C.c_int32, vp]
        L.oracle_recon.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32, vp, vp]
        L.oracle_rhothil_sum.argtypes = [vp, vp]
        L.oracle_rayf_profiles.argtypes = [vp, vp, vp]
        L.oracle_comp_alpha_press.argtypes = [vp, vp, vp, vp]
        L.oracle_set_turbulence.argtypes = [vp, vp, vp, vp, vp, vp]
        L.oracle_set_nudging.argtypes = [vp, vp, vp, vp]
        L.oracle_setup_tendencies.argtypes = [vp, vp]
        L.oracle_plant.argtypes = [vp, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Same call surface as olam_amd.gpu.Context, backed by the CPU restatement."""

    def __init__(self, grid, params, ref_types=False):
        self.grid = grid
        self.params = params
        self.h = C.c_void_p()
        rc = lib().oracle_create(C.byref(grid.desc), C.byref(params), int(ref_types), C.byref(self.h))
        if rc:
            raise RuntimeError(f"oracle_create failed ({rc})")

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def upload_state(self, st):
        d = st.desc()
        lib().oracle_upload_state(self.h, C.byref(d))

    def download_state(self, st):
        d = st.desc()
        lib().oracle_download_state(self.h, C.byref(d))

    def upload_scalar(self, n, q, t=None):
        q = np.ascontiguousarray(q, dtype=np.float32)
        t = None if t is None else np.ascontiguousarray(t, dtype=np.float32)
        rc = lib().oracle_upload_scalar(self.h, n, _p(q), _p(t))
        assert rc == 0

    def download_scalar(self, n):
        out = np.zeros((self.grid.mwa, self.grid.mza), np.float32)
        lib().oracle_download_scalar(self.h, n, _p(out))
        return out

    def upload_scalar_adj(self, n, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        lib().oracle_upload_scalar_adj(self.h, n, _p(lam))

    def download_scalar_adj(self, n):
        out = np.zeros((self.grid.mwa, self.grid.mza), np.float32)
        lib().oracle_download_scalar_adj(self.h, n, _p(out))
        return out

    def init_wrtv_rk(self):
        lib().oracle_init_wrtv_rk(self.h)

    def comp_alpha_press(self, rr_v, rr_w, theta):
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (rr_v, rr_w, theta)]
        lib().oracle_comp_alpha_press(self.h, *(_p(x) for x in a))

    def set_nudging(self, wmanud=None, vmanud=None, rhot_nud=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=dt)
             for x, dt in ((wmanud, np.float32), (vmanud, np.float32), (rhot_nud, np.float64))]
        lib().oracle_set_nudging(self.h, *(_p(x) for x in a))

    def set_turbulence(self, vkh, vkm, khtop, kmtop, akm_sfc):
        f = [np.ascontiguousarray(x, dtype=np.float32) for x in (vkh, vkm, akm_sfc)]
        i = [np.ascontiguousarray(x, dtype=np.int32) for x in (khtop, kmtop)]
        rc = lib().oracle_set_turbulence(self.h, _p(f[0]), _p(f[1]), _p(i[0]), _p(i[1]), _p(f[2]))
        if rc:
            raise RuntimeError(f"oracle set_turbulence failed ({rc})")

    def rayf_profiles(self):
        """Rayleigh friction profiles per level k = 1..mza: dict of rayf_cof, rayf_cofw,
        rayf_cofmix, b20 (float32 arrays) and the four k*_bot levels."""
        mza = self.grid.mza
        out = np.zeros(4 * mza, np.float32)
        kb = np.zeros(4, np.int32)
        lib().oracle_rayf_profiles(self.h, _p(out), _p(kb))
        d = {n: out[i * mza:(i + 1) * mza].copy() for i, n in enumerate(("rayf_cof", "rayf_cofw", "rayf_cofmix", "b20"))}
        d.update(zip(("krayf_bot", "krayfw_bot", "krayfdiv_bot", "krayfmix_bot"), (int(x) for x in kb)))
        return d

    def setup_tendencies(self):
        """Runs the setup part of prog_wrtv_rk; returns thilt_short, vmxet_short,
        vmyet_short, vmzet_short ((mwa, mza) each) and vmt_short ((mva, mza)), float64."""
        g = self.grid
        nw, nv = g.mwa * g.mza, g.mva * g.mza
        out = np.zeros(4 * nw + nv)
        rc = lib().oracle_setup_tendencies(self.h, _p(out))
        if rc:
            raise RuntimeError(f"oracle setup_tendencies failed ({rc})")
        names = ("thilt_s", "vmxet_s", "vmyet_s", "vmzet_s")
        d = {n: out[i * nw:(i + 1) * nw].reshape(g.mwa, g.mza) for i, n in enumerate(names)}
        d["vmt_s"] = out[4 * nw:].reshape(g.mva, g.mza)
        return d

    def tend0(self):
        lib().oracle_tend0(self.h)

    def prog_wrtv_rk(self):
        rc = lib().oracle_prog_wrtv_rk(self.h)
        if rc:
            raise RuntimeError(f"oracle prog_wrtv_rk failed ({rc})")

    def timeavg_momsc(self):
        lib().oracle_timeavg_momsc(self.h)

    def scalar_transport_rk(self, rho_old=None):
        r = None if rho_old is None else np.ascontiguousarray(rho_old, dtype=np.float32)
        lib().oracle_scalar_transport_rk(self.h, _p(r))

    def scalar_transport_rk_adj(self):
        rc = lib().oracle_scalar_transport_rk_adj(self.h)
        if rc:
            raise RuntimeError("oracle adjoint failed")

    def mass_sums(self):
        out = np.zeros(1 + self.params.num_scalar)
        lib().oracle_mass_sums(self.h, _p(out))
        return out

    def tracer_maxmin(self):
        """(num_scalar, 2) float32: max and min of each tracer over the prognostic levels."""
        out = np.zeros(2 * self.params.num_scalar, np.float32)
        lib().oracle_tracer_maxmin(self.h, _p(out))
        return out.reshape(-1, 2)

    def cfl_report(self):
        out = np.zeros(4)
        lib().oracle_cfl_report(self.h, _p(out))
        return out

    def set_scalar_pdef(self, n, pdef):
        assert lib().oracle_set_scalar_pdef(self.h, int(n), int(bool(pdef))) == 0

    def set_scalar_sgsmix(self, n, on):
        assert lib().oracle_set_scalar_sgsmix(self.h, int(n), int(bool(on))) == 0

    def tape_scales(self, n, which):
        """Test hook: taped limiter scales of tracer n, which 0 = faces (mva, mza), 1 = W levels."""
        rows = self.grid.mva if which == 0 else self.grid.mwa
        out = np.zeros((rows, self.grid.mza))
        assert lib().oracle_tape_scales(self.h, int(n), int(which), _p(out)) == 0
        return out

    def recon(self, field, order, bounded_coef=True, bounded_off=True):
        """Test hook: reconstruction coefficients (5, mwa, mza) of a (mwa, mza) field and
        its face values (2, mva, mza) on the iw1 / iw2 sides of every face."""
        g = self.grid
        f = np.ascontiguousarray(field, dtype=np.float32)
        coefs = np.zeros((5, g.mwa, g.mza))
        faces = np.zeros((2, g.mva, g.mza))
        rc = lib().oracle_recon(self.h, _p(f), int(order), int(bool(bounded_coef)), int(bool(bounded_off)),
                                _p(coefs), _p(faces))
        if rc != 0:
            raise ValueError(f"oracle_recon failed ({rc})")
        return coefs, faces

    def rhothil_sum(self):
        """Test hook: sum of rho * thil * volt over prognostic cells, full precision."""
        out = C.c_double()
        lib().oracle_rhothil_sum(self.h, C.byref(out))
        return out.value

    def plant(self, code):
        """Test hook: plant a deliberate mistake in one term (codes at oracle_plant in
        olam_oracle.cpp), so a pin can show that it catches it."""
        lib().oracle_plant(self.h, int(code))

    def set_frozen_limiter(self, on):
        """Test hook: later forward steps apply the limiter scales of the last step."""
        lib().oracle_set_frozen_limiter(self.h, int(bool(on)))


def tridiffo(a, b, c, d):
    """The oracle's Thomas solve of a x(i-1) + b x(i) + c x(i+1) = d."""
    a, b, c, d = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d))
    x = np.zeros_like(d)
    lib().oracle_tridiffo(len(d), _p(a), _p(b), _p(c), _p(d), _p(x))
    return x


def set_threads(n):
    return lib().oracle_set_threads(int(n))
