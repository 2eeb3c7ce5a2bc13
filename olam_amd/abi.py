"""ctypes mirror of include/olamgpu.h and conversion of meshes/states to Fortran layout.

The C-ABI takes the reference's Fortran layout: column-major (mza, n) arrays with
the vertical index contiguous and a dummy horizontal index 1, and neighbour tables
holding 1-based Fortran indices.  In numpy that is a C-ordered array of shape
(n, mza) whose row 0 is the dummy point; real point i (0-based) is row i + 1 and
Fortran index i + 2.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import mesh as M

PREC_F32, PREC_MIXED, PREC_F64 = 0, 1, 2
PREC_NAMES = {"f32": PREC_F32, "mixed": PREC_MIXED, "f64": PREC_F64}

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("mza", C.c_int32), ("mwa", C.c_int32), ("mva", C.c_int32), ("mma", C.c_int32),
        ("w_npoly", _ip), ("w_iv", _ip), ("w_iw", _ip), ("w_im", _ip), ("w_dirv", _fp),
        ("w_ecvec_vx", _fp), ("w_ecvec_vy", _fp), ("w_ecvec_vz", _fp),
        ("gxps_coef", _fp), ("gyps_coef", _fp), ("xx_yy", _fp),
        ("v_iw", _ip), ("v_im", _ip), ("v_dxps", _fp), ("v_dyps", _fp),
        ("m_iv", _ip), ("m_iw", _ip), ("m_im", _ip), ("xx_yy_m", _fp),
        ("lpw", _ip), ("lpv", _ip), ("lpm", _ip),
        ("wnx", _fp), ("wny", _fp), ("wnz", _fp), ("arw0", _fp),
        ("vnx", _fp), ("vny", _fp), ("vnz", _fp),
        ("dnu", _fp), ("dniu", _fp), ("dnv", _fp), ("dniv", _fp),
        ("arm0", _fp),
        ("dzim", _fp), ("dzit", _fp), ("zfacit", _fp),
        ("zwgt_top8", _dp), ("zwgt_bot8", _dp), ("gdz_wgtm8", _dp), ("gdz_wgtp8", _dp),
        ("a_w", _fp), ("a_wu", _fp),
        ("arw", _fp), ("volti", _fp), ("volwi", _fp), ("volt", _dp),
        ("arv", _fp), ("volvi", _fp),
        ("xy_h", _fp), ("xy_hu", _fp),
        ("xx0_v", _fp), ("xy0_v", _fp), ("yy0_v", _fp),
        ("xx0_vu", _fp), ("xy0_vu", _fp), ("yy0_vu", _fp),
        ("zm", _fp), ("zt", _fp), ("th01d", _fp), ("dn01d", _fp), ("dzit_bot", _fp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("num_scalar", C.c_int32), ("nrk_wrtv", C.c_int32),
        ("nrk_scal", C.c_int32), ("thil_horiz_adv_order", C.c_int32),
        ("wind_horiz_adv_order", C.c_int32), ("scal_horiz_adv_order", C.c_int32),
        ("iscal_monot", C.c_int32), ("divh_damp_level", C.c_int32),
        ("vort_damp_level", C.c_int32), ("divh_damp_fact", C.c_float),
        ("vort_damp_fact", C.c_float), ("dtsm", C.c_double), ("dtlm", C.c_double),
        ("nacoust", C.c_int32), ("adjoint_tape", C.c_int32), ("fcoriol", C.c_double),
        ("ithil_monot", C.c_int32), ("iwind_monot", C.c_int32),
        ("initial", C.c_int32),
        ("rayf_zmin", C.c_float), ("rayf_fact", C.c_float), ("rayf_expon", C.c_float),
        ("rayfw_zmin", C.c_float), ("rayfw_fact", C.c_float), ("rayfw_expon", C.c_float),
        ("rayfdiv_zmin", C.c_float), ("rayfdiv_fact", C.c_float), ("rayfdiv_expon", C.c_float),
        ("rayfmix_zmin", C.c_float), ("rayfmix_fact", C.c_float), ("rayfmix_expon", C.c_float),
        ("akmin", C.c_float),
        ("nudflag", C.c_int32), ("nud_preserve_total_mass", C.c_int32),
        ("nud_preserve_mix_ratio", C.c_int32),
    ]


STATE_FIELDS = [
    ("thil", "w", np.float32), ("wmc", "w", np.float32), ("wc", "w", np.float32),
    ("vxe", "w", np.float32), ("vye", "w", np.float32), ("vze", "w", np.float32),
    ("rho", "w", np.float64), ("press", "w", np.float64),
    ("vmc", "v", np.float32), ("vc", "v", np.float32),
    ("wmasc", "w", np.float32), ("vmasc", "v", np.float32),
    ("alpha_press", "w", np.float32), ("pwfac", "w", np.float32), ("pvfac", "v", np.float32),
    ("thilt", "w", np.float32), ("vmxet", "w", np.float32), ("vmyet", "w", np.float32),
    ("vmzet", "w", np.float32), ("vmt", "v", np.float32),
]


class StateDesc(C.Structure):
    _fields_ = [(name, _dp if dt == np.float64 else _fp) for name, _, dt in STATE_FIELDS]


RAYF_KINDS = ("rayf", "rayfw", "rayfdiv", "rayfmix")


def make_params(precision="f32", num_scalar=1, nrk_wrtv=3, nrk_scal=3, order=2, iscal_monot=0,
                divh_damp_level=3, vort_damp_level=3, divh_damp_fact=0.1, vort_damp_fact=0.075,
                dtsm=60.0, dtlm=None, nacoust=1, adjoint_tape=1, ithil_monot=0, iwind_monot=0,
                initial=1, rayf=None, akmin=0.0, nudflag=0, nud_preserve_total_mass=1,
                nud_preserve_mix_ratio=1):
    """Namelist defaults of oname_coms.f90:190-238, with RK3 for scalars.

    ``rayf`` maps a Rayleigh friction kind ("rayf", "rayfw", "rayfdiv", "rayfmix") to
    (zmin, fact, expon); kinds left out keep the namelist default (30 km, 0, 1: off).
    """
    if isinstance(precision, str):
        precision = PREC_NAMES[precision]
    if dtlm is None:
        dtlm = dtsm * nacoust
    rf = []
    for kind in RAYF_KINDS:
        rf.extend((rayf or {}).get(kind, (30000.0, 0.0, 1.0)))
    return Params(precision, num_scalar, nrk_wrtv, nrk_scal, order, order, order, iscal_monot,
                  divh_damp_level, vort_damp_level, divh_damp_fact, vort_damp_fact, dtsm, dtlm,
                  nacoust, adjoint_tape, 2.0 * M.OMEGA, ithil_monot, iwind_monot, initial, *rf, akmin,
                  nudflag, nud_preserve_total_mass, nud_preserve_mix_ratio)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class FortranGrid:
    """Mesh + vertical grid converted to the reference's Fortran layout.

    Holds every array it hands out so the pointers in ``desc`` stay valid.
    """

    def __init__(self, mesh: M.Mesh, vg: dict, cm: dict | None = None):
        if cm is None:
            cm = M.column_metrics(mesh, vg)
        self.mesh, self.vg, self.cm = mesh, vg, cm
        mza = vg["mza"]
        self.mza = mza
        self.mwa, self.mva, self.mma = mesh.nw + 1, mesh.nv + 1, mesh.nm + 1
        g = mesh.geo
        keep = {}

        def tab(a, width, fill=1):          # (n, width) 0-based index table -> (n+1, width) Fortran
            out = np.full((a.shape[0] + 1, width), fill, dtype=np.int32)
            out[1:, : a.shape[1]] = a + 2
            return out

        def per(a, dtype=np.float32, width=None):   # per-point array with dummy row 0
            if a.ndim == 1:
                out = np.zeros(a.shape[0] + 1, dtype=dtype)
                out[1:] = a
            else:
                out = np.zeros((a.shape[0] + 1, width or a.shape[1]), dtype=dtype)
                out[1:, : a.shape[1]] = a
            return out

        def lev(a, dtype=np.float32):      # per-level [k] (k=1..mza) -> (mza,)
            return np.ascontiguousarray(a[1 : mza + 1], dtype=dtype)

        def col(a, dtype=np.float32):      # (mza+1, n) [k, i] -> (n+1, mza) Fortran
            out = np.zeros((a.shape[1] + 1, mza), dtype=dtype)
            out[1:, :] = a[1 : mza + 1, :].T
            return out

        npoly = per(mesh.npoly.astype(np.int32), np.int32)
        keep["w_npoly"] = npoly
        act = np.arange(7)[None, :] < mesh.npoly[:, None]
        keep["w_iv"] = tab(np.where(act, mesh.w_iv, -1), 7)
        keep["w_iw"] = tab(np.where(act, mesh.w_iw, -1), 7)
        keep["w_im"] = tab(np.where(act, mesh.w_im, -1), 7)
        keep["w_dirv"] = per(mesh.dirv, np.float32)
        keep["w_ecvec_vx"] = per(g["ecvec_vx"])
        keep["w_ecvec_vy"] = per(g["ecvec_vy"])
        keep["w_ecvec_vz"] = per(g["ecvec_vz"])
        keep["gxps_coef"] = per(g["gxps_coef"])
        keep["gyps_coef"] = per(g["gyps_coef"])
        keep["xx_yy"] = per(g["xx_yy"])
        keep["v_iw"] = tab(mesh.v_iw, 2)
        keep["v_im"] = tab(mesh.v_im, 2)
        keep["v_dxps"] = per(g["dxps"])
        keep["v_dyps"] = per(g["dyps"])
        keep["m_iv"] = tab(mesh.m_iv, 3)
        keep["m_iw"] = tab(mesh.m_iw, 3)
        keep["m_im"] = tab(mesh.m_im, 3)
        keep["xx_yy_m"] = per(g["xx_yy_m"])
        keep["lpw"] = np.full(self.mwa, 2, np.int32)
        keep["lpv"] = np.full(self.mva, 2, np.int32)
        keep["lpm"] = np.full(self.mma, 2, np.int32)
        for name in ("wnx", "wny", "wnz", "arw0"):
            keep[name] = per(g[name])
        for name in ("vnx", "vny", "vnz", "dnu", "dniu", "dnv", "dniv"):
            keep[name] = per(g[name])
        keep["arm0"] = per(g["arm0"])
        for name in ("dzim", "dzit", "zfacit"):
            keep[name] = lev(vg[name])
        for name in ("zwgt_top8", "zwgt_bot8", "gdz_wgtm8", "gdz_wgtp8"):
            keep[name] = lev(vg[name], np.float64)
        # a_w(mza,2,2) Fortran: offset (k-1) + mza*((i-1) + 2*(j-1)) -> numpy [j, i, k]
        for name in ("a_w", "a_wu"):
            a = vg[name][1 : mza + 1]                      # [k, i, j]
            keep[name] = np.ascontiguousarray(a.transpose(2, 1, 0), dtype=np.float32)
        keep["arw"] = col(cm["arw"])
        keep["volti"] = col(cm["volti"])
        keep["volwi"] = col(cm["volwi"])
        keep["volt"] = col(cm["volt"], np.float64)
        keep["arv"] = col(cm["arv"])
        keep["volvi"] = col(cm["volvi"])
        # xy_h(5,7,mwa) Fortran -> numpy [iw, n, term]; xx0_v(2,mva) -> [iv, side]
        for name in ("xy_h", "xy_hu"):
            keep[name] = per(g[name].reshape(mesh.nw, 35)).reshape(self.mwa, 7, 5)
        for name in ("xx0_v", "xy0_v", "yy0_v", "xx0_vu", "xy0_vu", "yy0_vu"):
            keep[name] = per(g[name])
        keep["zm"] = lev(vg["zm"])
        keep["zt"] = lev(vg["zt"])
        keep["dzit_bot"] = lev(1.0 / np.where(vg["dzt_bot"] != 0, vg["dzt_bot"], 1.0))
        keep["th01d"] = keep["dn01d"] = None     # set_sounding()
        self.arrays = keep
        d = MeshDesc()
        d.mza, d.mwa, d.mva, d.mma = mza, self.mwa, self.mva, self.mma
        for fname, ftype in MeshDesc._fields_[4:]:
            a = keep[fname]
            if a is not None:
                setattr(d, fname, _ptr(a, ftype._type_))
        self.desc = d
        self.pwfac = col(np.repeat(vg["dzim"][:, None], mesh.nw, axis=1))
        self.pvfac = col(cm["pvfac"])

    def set_sounding(self, th01d, dn01d):
        """The 1-D reference sounding (misc_coms.f90:112-113: potential temperature and
        density per level k = 1..mza) that Rayleigh friction relaxes thil toward."""
        for name, a in (("th01d", th01d), ("dn01d", dn01d)):
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.shape != (self.mza,):
                raise ValueError(f"{name} must have shape ({self.mza},)")
            self.arrays[name] = a
            setattr(self.desc, name, _ptr(a, C.c_float))

    # helpers to move fields between (n, mza) real-point arrays and Fortran layout
    def to_f(self, a, kind="w", dtype=np.float32):
        n = self.mwa if kind == "w" else (self.mva if kind == "v" else self.mma)
        out = np.zeros((n, self.mza), dtype=dtype)
        out[1:] = a
        return out

    @staticmethod
    def from_f(a):
        return a[1:]


class FortranState:
    """Host buffers for olamgpu_state in Fortran layout (numpy (n, mza))."""

    def __init__(self, grid: FortranGrid, fields: dict | None = None):
        self.grid = grid
        self.arrays = {}
        for name, kind, dt in STATE_FIELDS:
            n = grid.mwa if kind == "w" else grid.mva
            self.arrays[name] = np.zeros((n, grid.mza), dtype=dt)
        self.arrays["alpha_press"][:] = M.ALPHA_DRY
        self.arrays["pwfac"][:] = grid.pwfac
        self.arrays["pvfac"][:] = grid.pvfac
        if fields:
            for k, v in fields.items():
                self.arrays[k][...] = v

    def desc(self, only=None):
        d = StateDesc()
        for name, kind, dt in STATE_FIELDS:
            if only is not None and name not in only:
                continue
            a = self.arrays[name]
            n = self.grid.mva if kind == "v" else self.grid.mwa
            if a.dtype != dt or a.shape != (n, self.grid.mza) or not a.flags["C_CONTIGUOUS"]:
                raise ValueError(f"state member {name} must be a contiguous {np.dtype(dt).name} "
                                 f"array of shape {(n, self.grid.mza)}")
            setattr(d, name, _ptr(a, C.c_double if dt == np.float64 else C.c_float))
        return d

    def copy(self):
        s = FortranState(self.grid)
        for k, v in self.arrays.items():
            s.arrays[k][...] = v
        return s
