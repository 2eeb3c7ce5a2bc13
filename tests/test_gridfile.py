"""OLAM gridfile and history I/O (olam_amd/gridfile.py; the reference's gridfile_write
omodel/olam_grid.f90:828-1450 and history_write omodel/history_write.f90:1-190), on the
CPU:

* a grid written and read back gives bit-identical device inputs (every array of the
  C-ABI's mesh descriptor), and an oracle step on the re-read grid is bit-identical;
* the datasets carry the reference's names, shapes (Fortran arrays with the dimensions
  reversed) and types;
* without the derived datasets (an OLAM-written gridfile has none) the reader derives
  them with the reference's formulas: gxps_coef from gxps1/gxps2 (mem_grid.f90:581-590)
  to FP32 rounding, inverse volumes, vertical weights and Laplacian weights;
* a history file round-trips the state and the tracers exactly.
"""

import os

import numpy as np
import pytest

from olam_amd import abi, cases, gridfile as GF, mesh as M
from oracle import Oracle

pytestmark = pytest.mark.skipif(not os.path.exists(GF.LIB), reason="libolamio.so not built (no HDF5)")


@pytest.fixture(scope="module")
def grid():
    return abi.FortranGrid(M.build_mesh(6), M.vertical_grid(12))


def test_gridfile_round_trip_bit_identical(grid, tmp_path):
    p = str(tmp_path / "grid.h5")
    GF.write_gridfile(p, grid)
    g2 = GF.read_gridfile(p)
    assert (g2.mza, g2.mwa, g2.mva, g2.mma) == (grid.mza, grid.mwa, grid.mva, grid.mma)
    for name, a in grid.arrays.items():
        b = g2.arrays[name]
        if a is None:
            assert b is None, name
            continue
        assert a.dtype == b.dtype and a.shape == b.shape, name
        assert np.array_equal(a, b), name
    assert np.array_equal(grid.pwfac, g2.pwfac) and np.array_equal(grid.pvfac, g2.pvfac)
    # the same step on both grids
    outs = []
    for gg in (grid, g2):
        st = cases.jet_state(grid)
        st2 = abi.FortranState(gg)
        for k, v in st.arrays.items():
            st2.arrays[k][...] = v
        o = Oracle(gg, abi.make_params("f32", num_scalar=1, dtsm=120.0))
        o.upload_state(st2)
        o.upload_scalar(1, cases.cosine_bell(grid))
        o.tend0()
        o.prog_wrtv_rk()
        o.timeavg_momsc()
        o.scalar_transport_rk()
        res = abi.FortranState(gg)
        o.download_state(res)
        outs.append((res.arrays["thil"], res.arrays["vmc"], o.download_scalar(1)))
        o.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_gridfile_names_shapes_types(grid, tmp_path):
    p = str(tmp_path / "grid.h5")
    GF.write_gridfile(p, grid)
    with GF.H5File(p) as f:
        assert f.scalar("NZA") == grid.mza and f.scalar("NWA") == grid.mwa
        for name, shape, dt in (("ZT", (grid.mza,), np.float32), ("itab_w%iv", (grid.mwa, 7), np.int32),
                                ("itab_v%iw", (grid.mva, 2), np.int32), ("itab_m%iv", (grid.mma, 3), np.int32),
                                ("ARW", (grid.mwa, grid.mza), np.float32),
                                ("VOLT", (grid.mwa, grid.mza), np.float64),
                                ("ARV", (grid.mva, grid.mza), np.float32),
                                ("itab_w%gxps1", (grid.mwa, 7), np.float32), ("DNV", (grid.mva,), np.float32)):
            a = f.read(name)
            assert a.shape == shape and a.dtype == dt, name
        # 1-based Fortran indices with the dummy point 1
        iw = f.read("itab_v%iw")
        assert iw[1:].min() >= 2 and iw[1:].max() <= grid.mwa


def test_gridfile_derives_what_olam_derives(grid, tmp_path):
    """An OLAM gridfile carries gxps1/gxps2 but not gxps_coef, nor the inverse volumes or
    the vertical weights: drop the derived datasets and read again."""
    full = str(tmp_path / "grid.h5")
    GF.write_gridfile(full, grid)
    bare = str(tmp_path / "bare.h5")
    derived = {"GXPS_COEF", "GYPS_COEF", "XX_YY", "XX_YY_M", "VOLTI", "VOLWI", "VOLVI", "PVFAC",
               "DZT_TOP", "DZT_BOT", "ZWGT_TOP8", "ZWGT_BOT8", "GDZ_WGTM8", "GDZ_WGTP8"}
    names = ["NZP", "NXP", "MDOMAIN", "NGRIDS", "NZA", "NMA", "NVA", "NWA", "MRLS", "ZM", "ZT", "DZM", "DZT",
             "DZIM", "DZIT", "ZFACM", "ZFACT", "ZFACIT", "ZFACM2", "ZFACIM", "ZFACIM2", "LPM", "XEM", "YEM",
             "ZEM", "ARM0", "itab_m%npoly", "itab_m%iv", "itab_m%iw", "itab_m%im", "LPV", "XEV", "YEV", "ZEV",
             "UNX", "UNY", "UNZ", "DNU", "DNV", "DNIU", "DNIV", "VNX", "VNY", "VNZ", "itab_v%iw", "itab_v%im",
             "itab_v%dxps", "itab_v%dyps", "LPW", "XEW", "YEW", "ZEW", "GLATW", "GLONW", "ARW0", "WNX", "WNY",
             "WNZ", "itab_w%npoly", "itab_w%iv", "itab_w%iw", "itab_w%im", "itab_w%dirv", "itab_w%farm",
             "itab_w%farv", "itab_w%gxps1", "itab_w%gyps1", "itab_w%gxps2", "itab_w%gyps2",
             "itab_w%ecvec_vx", "itab_w%ecvec_vy", "itab_w%ecvec_vz", "ARV", "ARW", "VOLT", "A_W", "A_WU"]
    with GF.H5File(full) as src, GF.H5File(bare, "w") as dst:
        for n in names:
            assert n not in derived
            dst.write(n, src.read(n))
    g2 = GF.read_gridfile(bare)
    A, B = grid.arrays, g2.arrays
    for name in ("gxps_coef", "gyps_coef", "xx_yy", "xx_yy_m", "volti", "volwi", "volvi"):
        a, b = A[name].astype(np.float64), B[name].astype(np.float64)
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err < 2e-6, (name, err)
    for name in ("zwgt_top8", "zwgt_bot8", "gdz_wgtm8", "gdz_wgtp8", "dzit_bot"):
        assert np.allclose(A[name], B[name], rtol=1e-6, atol=0), name
    assert np.allclose(grid.pvfac, g2.pvfac, rtol=1e-6, atol=0)


def test_history_round_trip(grid, tmp_path):
    st = cases.jet_state(grid)
    tr = {"CO2": cases.cosine_bell(grid), "Q2": cases.lognormal_tracers(grid, 1)[0]}
    p = str(tmp_path / "hist.h5")
    GF.write_history(p, st, tr)
    st2, tr2 = GF.read_history(p, grid, tracer_names=("CO2", "Q2"))
    for k in GF.HISTORY_FIELDS:
        assert np.array_equal(st.arrays[k], st2.arrays[k]), k
        assert st.arrays[k].dtype == st2.arrays[k].dtype
    for k in tr:
        assert np.array_equal(tr[k], tr2[k])
    with GF.H5File(p) as f:
        assert f.read("RHO").dtype == np.float64 and f.read("THIL").dtype == np.float32
