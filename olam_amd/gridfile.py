"""OLAM gridfiles and history files in HDF5.

The reference writes its grid once (gridfile_write, omodel/olam_grid.f90:828-1450) and
reads it back at every start (gridfile_read, :1939-2340), and writes the model state
as history files (history_write, omodel/history_write.f90:1-190) under the names its
variable tables give the fields (mem_basic.f90:150-185: VMC, VC, WMC, WC, THIL, RHO,
PRESS, VXE, VYE, VZE; tracers under their scalar_tab names).  This module reads and
writes both with the same dataset names, shapes and types, so the GPU path can run from
an OLAM grid and an OLAM state:

* every array is stored as its Fortran array with the dummy point 1, which the Fortran
  HDF5 interface lays out in C order with the dimensions reversed: a (mza, mwa) field
  is a (mwa, mza) dataset, exactly the numpy layout of ``abi.FortranGrid`` and
  ``abi.FortranState``; neighbour tables hold 1-based Fortran indices;
* REAL is float32, real(r8) float64, INTEGER int32, as in the reference.

OLAM's gridfile holds the grid geometry and tables (itab_w, itab_v, itab_m members,
the metric arrays ARW, ARV, VOLT, the level heights); the model derives the rest when
it starts (gxps_coef from gxps1/gxps2, mem_grid.f90:581-590; volti, volwi, volvi; the
vertical weights of alloc_gridz_other, mem_grid.f90:350-440; the Laplacian and
reconstruction weights of alloc_adv, mem_adv.f90:23-446).  ``write_gridfile`` adds those
derived arrays as extra datasets (upper-case names OLAM does not use), so a file
written here reads back into bit-identical device inputs; ``read_gridfile`` derives any
of them that a file lacks, with the reference's formulas.

The HDF5 calls go through ``olam_amd/lib/libolamio.so`` (olam_amd/csrc/olamio.cpp,
linked against the image's C libhdf5).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi, mesh as M

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(PKG, "lib", "libolamio.so")
SRC = os.path.join(PKG, "csrc", "olamio.cpp")
HDF5_PREFIX = "/opt/conda"

_T = {np.dtype(np.int32): 0, np.dtype(np.float32): 1, np.dtype(np.float64): 2}
_D = {0: np.int32, 1: np.float32, 2: np.float64}


def build(verbose=False):
    """Compile the HDF5 storage layer (g++ against the image's C libhdf5)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    import subprocess
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", f"-I{HDF5_PREFIX}/include", SRC,
           f"-L{HDF5_PREFIX}/lib", "-lhdf5", f"-Wl,-rpath,{HDF5_PREFIX}/lib", "-o", LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (olam_amd.gridfile.build())")
        L = C.CDLL(LIB)
        L.olamio_create.argtypes = [C.c_char_p]
        L.olamio_create.restype = C.c_int64
        L.olamio_open.argtypes = [C.c_char_p, C.c_int]
        L.olamio_open.restype = C.c_int64
        L.olamio_close.argtypes = [C.c_int64]
        L.olamio_write.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_void_p]
        L.olamio_exists.argtypes = [C.c_int64, C.c_char_p]
        L.olamio_info.argtypes = [C.c_int64, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.olamio_read.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p]
        _lib = L
    return _lib


class H5File:
    """An HDF5 file of named datasets (root group), opened for writing ("w", truncates)
    or reading ("r")."""

    def __init__(self, path, mode="r"):
        L = lib()
        self.path = path
        self.h = L.olamio_create(path.encode()) if mode == "w" else L.olamio_open(path.encode(), 0)
        if self.h < 0:
            raise OSError(f"cannot open {path} ({mode})")

    def close(self):
        if self.h >= 0:
            lib().olamio_close(self.h)
            self.h = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write(self, name, a):
        a = np.ascontiguousarray(a)
        if a.dtype not in _T:
            a = a.astype(np.float64 if a.dtype.kind == "f" else np.int32)
        if a.ndim == 0:
            a = a.reshape(1)
        dims = (C.c_int64 * a.ndim)(*a.shape)
        rc = lib().olamio_write(self.h, name.encode(), _T[a.dtype], a.ndim, dims, a.ctypes.data)
        if rc != 0:
            raise OSError(f"cannot write dataset {name} to {self.path}")

    def has(self, name):
        return lib().olamio_exists(self.h, name.encode()) == 1

    def read(self, name, dtype=None):
        t = C.c_int()
        dims = (C.c_int64 * 4)()
        rank = lib().olamio_info(self.h, name.encode(), C.byref(t), dims)
        if rank < 0:
            raise KeyError(f"{name} not in {self.path}")
        dt = np.dtype(dtype) if dtype is not None else np.dtype(_D[t.value])
        out = np.empty(tuple(dims[i] for i in range(rank)), dtype=dt)
        if lib().olamio_read(self.h, name.encode(), _T[dt], out.ctypes.data) != 0:
            raise OSError(f"cannot read dataset {name} from {self.path}")
        return out

    def scalar(self, name):
        return self.read(name).reshape(-1)[0]


# ---------------------------------------------------------------- gridfile

def _per(a, dtype):
    """Per-point array with the dummy point 1 in front (index 0 here)."""
    a = np.asarray(a)
    out = np.zeros((a.shape[0] + 1,) + a.shape[1:], dtype=dtype)
    out[1:] = a
    return out


def _tab(a, width):
    """0-based (n, <=width) index table -> 1-based Fortran (n + 1, width), dummy 1."""
    out = np.ones((a.shape[0] + 1, width), np.int32)
    out[1:, : a.shape[1]] = a + 2
    return out


def _col(cm_arr, mza, dtype=np.float32):
    """[k, i] (mza + 1, n) metric -> Fortran (n + 1, mza)."""
    out = np.zeros((cm_arr.shape[1] + 1, mza), dtype=dtype)
    out[1:] = cm_arr[1 : mza + 1].T
    return out


def _uncol(f, mza):
    """Fortran (n + 1, mza) -> [k, i] (mza + 1, n), row 0 unused."""
    out = np.zeros((mza + 1, f.shape[0] - 1), dtype=np.float64)
    out[1:] = f[1:].T
    return out


# per-level arrays of the gridfile (olam_grid.f90:938-951) and derived ones added here
_LEVELS = ("ZM", "ZT", "DZM", "DZT", "DZIM", "DZIT", "ZFACM", "ZFACT", "ZFACIT", "ZFACM2")
_LEVELS_DERIVED = ("DZT_TOP", "DZT_BOT", "ZWGT_TOP8", "ZWGT_BOT8", "GDZ_WGTM8", "GDZ_WGTP8")


def write_gridfile(path, grid: abi.FortranGrid):
    """The grid of ``grid`` as an OLAM gridfile (gridfile_write's dataset names, shapes
    and types) plus the derived arrays the model computes at start, so the file reads
    back into bit-identical device inputs."""
    mesh, vg, cm = grid.mesh, grid.vg, grid.cm
    g = mesh.geo
    mza = grid.mza
    A = grid.arrays
    act = np.arange(7)[None, :] < mesh.npoly[:, None]
    with H5File(path, "w") as f:
        for name, v in (("NZP", mza - 1), ("NXP", mesh.nxp), ("MDOMAIN", 0), ("NGRIDS", 1), ("NZA", mza),
                        ("NMA", grid.mma), ("NVA", grid.mva), ("NWA", grid.mwa), ("MRLS", 1)):
            f.write(name, np.array([v], np.int32))
        for name in _LEVELS:
            f.write(name, np.asarray(vg[name.lower()][1 : mza + 1], np.float32))
        f.write("ZFACIM", np.asarray(1.0 / vg["zfacm"][1 : mza + 1], np.float32))
        f.write("ZFACIM2", np.asarray(1.0 / vg["zfacm2"][1 : mza + 1], np.float32))
        # M points (olam_grid.f90:955-964, 1017-1086)
        f.write("LPM", A["lpm"])
        for name in ("xem", "yem", "zem"):
            f.write(name.upper(), _per(g[name], np.float32))
        f.write("ARM0", A["arm0"])
        f.write("itab_m%npoly", np.full(grid.mma, 3, np.int32))
        for name in ("iv", "iw", "im"):
            f.write(f"itab_m%{name}", A[f"m_{name}"])
        # V points (:967-986, 1090-1180)
        f.write("LPV", A["lpv"])
        for name in ("xev", "yev", "zev", "unx", "uny", "unz"):
            f.write(name.upper(), _per(g[name], np.float32))
        for name in ("dnu", "dnv", "dniu", "dniv", "vnx", "vny", "vnz"):
            f.write(name.upper(), A[name])
        f.write("itab_v%iw", A["v_iw"])
        f.write("itab_v%im", A["v_im"])
        f.write("itab_v%dxps", A["v_dxps"])
        f.write("itab_v%dyps", A["v_dyps"])
        # W points (:988-1003, 1190-1345)
        f.write("LPW", A["lpw"])
        for name in ("xew", "yew", "zew", "glatw", "glonw"):
            f.write(name.upper(), _per(g[name], np.float32))
        for name in ("arw0", "wnx", "wny", "wnz"):
            f.write(name.upper(), A[name])
        f.write("itab_w%npoly", A["w_npoly"])
        for name in ("iv", "iw", "im"):
            f.write(f"itab_w%{name}", A[f"w_{name}"])
        f.write("itab_w%dirv", A["w_dirv"])
        for name in ("farm", "farv", "gxps1", "gyps1", "gxps2", "gyps2"):
            f.write(f"itab_w%{name}", _per(np.where(act, g[name], 0.0), np.float32))
        for c in "xyz":
            f.write(f"itab_w%ecvec_v{c}", A[f"w_ecvec_v{c}"])
        # 3-D metrics (:1005-1015)
        f.write("ARV", A["arv"])
        f.write("ARW", A["arw"])
        f.write("VOLT", A["volt"])
        # derived at OLAM's start (mem_grid.f90:350-618, mem_adv.f90:23-446)
        for name in _LEVELS_DERIVED:
            f.write(name, np.asarray(vg[name.lower()][1 : mza + 1], np.float64))
        f.write("A_W", A["a_w"])
        f.write("A_WU", A["a_wu"])
        for name in ("gxps_coef", "gyps_coef", "xx_yy", "xx_yy_m", "volti", "volwi", "volvi"):
            f.write(name.upper(), A[name])
        f.write("PVFAC", grid.pvfac)
        if g.get("xy_h") is not None:
            for name in ("xy_h", "xy_hu", "xx0_v", "xy0_v", "yy0_v", "xx0_vu", "xy0_vu", "yy0_vu"):
                f.write(name.upper(), A[name])


def _derive_gxps_coef(npoly, gps1, gps2):
    """mem_grid.f90:581-590: the gradient weight of neighbour j is point 1 of triangle j
    plus point 2 of triangle j - 1 (cyclic), FP32 as the reference."""
    n = npoly.shape[0]
    J = np.arange(7)[None, :]
    act = J < npoly[:, None]
    jm1 = (J - 1) % npoly[:, None]
    out = np.where(act, gps1 + np.take_along_axis(gps2, jm1, 1), np.float32(0)).astype(np.float32)
    assert out.shape == (n, 7)
    return out


def read_gridfile(path) -> abi.FortranGrid:
    """An ``abi.FortranGrid`` from an OLAM gridfile.  Arrays the file lacks that the
    model derives at start (the gradient weights, inverse volumes, vertical weights,
    Laplacian weights, pvfac) are derived here with the reference's formulas."""
    with H5File(path, "r") as f:
        mza = int(f.scalar("NZA"))
        mwa, mva, mma = int(f.scalar("NWA")), int(f.scalar("NVA")), int(f.scalar("NMA"))
        nw, nv, nm = mwa - 1, mva - 1, mma - 1
        nxp = int(f.scalar("NXP")) if f.has("NXP") else 0
        npoly = f.read("itab_w%npoly", np.int32)[1:].astype(np.int64)
        act = np.arange(7)[None, :] < npoly[:, None]

        def tab0(name, width):
            t = f.read(name, np.int32)[1:, :width].astype(np.int64) - 2
            return np.where(t < 0, 0, t)

        mesh = M.Mesh(nxp=nxp, nw=nw, nv=nv, nm=nm, npoly=npoly,
                      w_iv=np.where(act, tab0("itab_w%iv", 7), 0), w_iw=np.where(act, tab0("itab_w%iw", 7), 0),
                      w_im=np.where(act, tab0("itab_w%im", 7), 0),
                      dirv=f.read("itab_w%dirv", np.float32)[1:].astype(np.float64),
                      v_iw=tab0("itab_v%iw", 2), v_im=tab0("itab_v%im", 2),
                      m_iw=tab0("itab_m%iw", 3), m_iv=tab0("itab_m%iv", 3), m_im=tab0("itab_m%im", 3))
        mesh.w_canon, mesh.v_canon, mesh.m_canon = np.arange(nw), np.arange(nv), np.arange(nm)
        mesh.v_orient = np.ones(nv)
        g = mesh.geo

        def pt(name, dtype=np.float32):
            return f.read(name, dtype)[1:].astype(np.float64)

        for name in ("xem", "yem", "zem"):
            g[name] = pt(name.upper())
        for name in ("xev", "yev", "zev", "unx", "uny", "unz", "dnu", "dnv", "dniu", "dniv", "vnx", "vny", "vnz"):
            g[name] = pt(name.upper())
        for name in ("xew", "yew", "zew", "glatw", "glonw", "arw0", "wnx", "wny", "wnz"):
            g[name] = pt(name.upper())
        g["arm0"] = pt("ARM0")
        g["dxps"] = pt("itab_v%dxps")
        g["dyps"] = pt("itab_v%dyps")
        for c in "xyz":
            g[f"ecvec_v{c}"] = pt(f"itab_w%ecvec_v{c}")
        for name in ("farm", "farv", "gxps1", "gyps1", "gxps2", "gyps2"):
            g[name] = pt(f"itab_w%{name}")
        for name, src in (("gxps_coef", ("gxps1", "gxps2")), ("gyps_coef", ("gyps1", "gyps2"))):
            if f.has(name.upper()):
                g[name] = pt(name.upper())
            else:
                g[name] = _derive_gxps_coef(npoly, f.read(f"itab_w%{src[0]}", np.float32)[1:],
                                            f.read(f"itab_w%{src[1]}", np.float32)[1:]).astype(np.float64)
        # Laplacian weights, mem_adv.f90:121-147
        if f.has("XX_YY"):
            g["xx_yy"] = pt("XX_YY")
        else:
            ivj = mesh.w_iv
            g["xx_yy"] = np.where(act, (1.0 / g["dnv"][ivj]) * g["dnu"][ivj] / g["arw0"][:, None], 0.0)
        g["xx_yy_m"] = pt("XX_YY_M") if f.has("XX_YY_M") else \
            g["dnv"][mesh.m_iv] * (1.0 / g["dnu"][mesh.m_iv]) / g["arm0"][:, None]
        if f.has("XY_H"):
            g["xy_h"] = f.read("XY_H", np.float32)[1:].astype(np.float64)
            g["xy_hu"] = f.read("XY_HU", np.float32)[1:].astype(np.float64)
            for name in ("xx0_v", "xy0_v", "yy0_v", "xx0_vu", "xy0_vu", "yy0_vu"):
                g[name] = pt(name.upper())
        else:
            # alloc_adv (mem_adv.f90:161-354): the least-squares fits of order 3 from the
            # M points of each W fan in the W's polar-stereographic plane
            xw = np.stack([g["xew"], g["yew"], g["zew"]], 1)
            xm = np.stack([g["xem"], g["yem"], g["zem"]], 1)
            raxis = np.sqrt(xw[:, 0] ** 2 + xw[:, 1] ** 2)
            big = raxis >= 1.0e2
            frame = [np.repeat(v[:, None], 7, 1) for v in (
                raxis / M.ERAD, xw[:, 2] / M.ERAD, np.where(big, xw[:, 0] / np.maximum(raxis, 1e2), 1.0),
                np.where(big, xw[:, 1] / np.maximum(raxis, 1e2), 0.0))]
            xmn, ymn = M._ps(xm[mesh.w_im] - xw[:, None, :], frame[0], frame[1], frame[2], frame[3])
            g.update(M._quadratic_fit(mesh, np.where(act, xmn, 0.0).astype(np.float32),
                                      np.where(act, ymn, 0.0).astype(np.float32), g["arw0"]))
        # vertical grid: level arrays indexed by Fortran k (index 0 unused)
        vg = {"mza": mza}

        def lev(name, dtype=np.float32):
            a = np.zeros(mza + 1)
            a[1:] = f.read(name, dtype)
            return a

        for name in _LEVELS:
            vg[name.lower()] = lev(name)
        vg["zfacit"] = lev("ZFACIT")
        if f.has("DZT_TOP"):
            for name in _LEVELS_DERIVED:
                vg[name.lower()] = lev(name, np.float64)
        else:   # mem_grid.f90:350-440 (alloc_gridz_other), constant gravity
            zm, zt, dzt, dzim = vg["zm"], vg["zt"], vg["dzt"], vg["dzim"]
            vg["dzt_top"] = np.zeros(mza + 1)
            vg["dzt_bot"] = np.zeros(mza + 1)
            vg["dzt_top"][1] = zm[1] - zt[1]
            vg["dzt_bot"][1] = dzt[1] - vg["dzt_top"][1]
            for k in range(2, mza + 1):
                vg["dzt_top"][k] = zm[k] - zt[k]
                vg["dzt_bot"][k] = zt[k] - zm[k - 1]
            zt8, zb8 = np.zeros(mza + 1), np.zeros(mza + 1)
            for k in range(1, mza):
                zt8[k] = vg["dzt_bot"][k + 1] * dzim[k]
                zb8[k] = vg["dzt_top"][k] * dzim[k]
            zt8[mza], zb8[mza] = zt8[mza - 1], zb8[mza - 1]
            vg["zwgt_top8"], vg["zwgt_bot8"] = zt8, zb8
            vg["gdz_wgtp8"], vg["gdz_wgtm8"] = zt8 * M.GRAV, zb8 * M.GRAV
        for name in ("a_w", "a_wu"):
            a = f.read(name.upper(), np.float32)          # [j, i, k]
            full = np.zeros((mza + 1, 2, 2))
            full[1:] = a.transpose(2, 1, 0)
            vg[name] = full
        # column metrics, [k, i] with row 0 unused
        cm = {"arw": _uncol(f.read("ARW", np.float32), mza), "arv": _uncol(f.read("ARV", np.float32), mza),
              "volt": _uncol(f.read("VOLT", np.float64), mza)}
        if f.has("VOLTI"):
            for name in ("volti", "volwi", "volvi"):
                cm[name] = _uncol(f.read(name.upper(), np.float32), mza)
        else:
            volt = cm["volt"]
            cm["volti"] = np.zeros_like(volt)
            cm["volti"][2:] = 1.0 / volt[2:]
            cm["volwi"] = np.zeros_like(volt)
            cm["volwi"][2:mza] = 1.0 / (volt[2:mza] + volt[3:mza + 1])
            i1, i2 = mesh.v_iw[:, 0], mesh.v_iw[:, 1]
            cm["volvi"] = np.zeros((mza + 1, nv))
            cm["volvi"][2:] = 1.0 / (volt[2:, i1] + volt[2:, i2])
        cm["pvfac"] = _uncol(f.read("PVFAC", np.float32), mza) if f.has("PVFAC") else \
            vg["zfacit"][:, None] * g["dniv"][None, :]
    return abi.FortranGrid(mesh, vg, cm)


# ---------------------------------------------------------------- history files

# mem_basic.f90:150-185: the names under which history_write stores the state
HISTORY_FIELDS = {"vmc": "VMC", "vc": "VC", "wmc": "WMC", "wc": "WC", "thil": "THIL", "rho": "RHO",
                  "press": "PRESS", "vxe": "VXE", "vye": "VYE", "vze": "VZE"}


def write_history(path, state: abi.FortranState, tracers=None):
    """The prognostic state (and tracers: a dict name -> (mwa, mza) float32 array, the
    scalar_tab names) as an OLAM history file."""
    with H5File(path, "w") as f:
        for k, name in HISTORY_FIELDS.items():
            f.write(name, state.arrays[k])
        for name, q in (tracers or {}).items():
            f.write(name, np.ascontiguousarray(q, dtype=np.float32))


def read_history(path, grid: abi.FortranGrid, tracer_names=()):
    """(FortranState, {name: tracer}) from an OLAM history file written on this grid."""
    st = abi.FortranState(grid)
    tr = {}
    with H5File(path, "r") as f:
        for k, name in HISTORY_FIELDS.items():
            a = f.read(name)
            if a.shape != st.arrays[k].shape:
                raise ValueError(f"{name}: shape {a.shape}, grid expects {st.arrays[k].shape}")
            st.arrays[k][...] = a
        for name in tracer_names:
            tr[name] = f.read(name, np.float32)
    return st, tr
