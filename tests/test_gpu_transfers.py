"""Host <-> device transfers of the C-ABI (olamgpu_upload_state / download_state,
upload_scalar / download_scalar, upload_scalar_adj / download_scalar_adj): host arrays in
the reference's Fortran layout (mza, n + 1) with the dummy column, moved through pinned
bounce buffers and repacked on the device to the renumbered, padded layout.

* Every state member comes back bit-identical (FP32 members; rho and press in FP64 in
  the modes that keep them in FP64, rounded to FP32 in the FP32 mode), the dummy column
  comes back zero, on a mesh whose V fields (NXP=160, L=50: 157 MB) span several bounce
  chunks and end in a partial one.
* NULL members are skipped on upload (the device keeps its values) and not written on
  download (the host keeps its values), as the Fortran shim relies on.
* Tracers, their long-step tendencies and the adjoint fields round-trip the same way.
"""

import numpy as np
import pytest

from olam_amd import abi
from olam_amd.gpu import Context

pytestmark = pytest.mark.gpu


def _random_state(grid, seed):
    rng = np.random.default_rng(seed)
    st = abi.FortranState(grid)
    for name, kind, dt in abi.STATE_FIELDS:
        a = st.arrays[name]
        a[...] = rng.standard_normal(a.shape).astype(dt)
        a[0] = 0                                      # the dummy column is never read
    return st


@pytest.mark.parametrize("prec", ["f32", "mixed", "f64"])
def test_state_round_trip_multi_chunk(grid_factory, prec):
    grid = grid_factory(160, 50)
    print(f"transfers {prec}: NXP=160 L=50, V field {grid.mva * grid.mza * 4 / 1e6:.0f} MB", flush=True)
    st = _random_state(grid, 11)
    g = Context(grid, abi.make_params(prec, num_scalar=2, dtsm=60.0), device=0)
    try:
        g.upload_state(st)
        back = abi.FortranState(grid)
        for a in back.arrays.values():
            a[...] = np.nan                           # every element must be written
        g.download_state(back)
        for name, kind, dt in abi.STATE_FIELDS:
            want = st.arrays[name]
            if dt == np.float64 and prec == "f32":
                want = want.astype(np.float32).astype(np.float64)
            got = back.arrays[name]
            assert np.array_equal(got[1:], want[1:]), name
            assert np.all(got[0] == 0), name
    finally:
        g.close()


def test_null_members_are_skipped(grid_factory):
    grid = grid_factory(24, 30)
    st = _random_state(grid, 3)
    st2 = _random_state(grid, 4)
    g = Context(grid, abi.make_params("f32", num_scalar=1, dtsm=60.0), device=0)
    try:
        g.upload_state(st)
        g.upload_state(st2, only=("thilt", "vmxet", "vmt"))    # the shim's tendency upload
        host = abi.FortranState(grid)
        sentinel = np.float32(12345.0)
        for a in host.arrays.values():
            a[...] = sentinel
        g.download_state(host, only=("thil", "rho", "vmt", "thilt"))
        assert np.array_equal(host.arrays["thil"][1:], st.arrays["thil"][1:])
        assert np.array_equal(host.arrays["rho"][1:],
                              st.arrays["rho"][1:].astype(np.float32).astype(np.float64))
        assert np.array_equal(host.arrays["thilt"][1:], st2.arrays["thilt"][1:])
        assert np.array_equal(host.arrays["vmt"][1:], st2.arrays["vmt"][1:])
        assert np.all(host.arrays["wc"] == sentinel) and np.all(host.arrays["vmc"] == sentinel)
        full = abi.FortranState(grid)
        g.download_state(full)
        for name in ("wmc", "vxe", "vmc", "pvfac", "vmyet"):          # untouched members
            assert np.array_equal(full.arrays[name][1:], st.arrays[name][1:]), name
        assert np.array_equal(full.arrays["vmxet"][1:], st2.arrays["vmxet"][1:])
    finally:
        g.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tracer_and_adjoint_round_trip(grid_factory, prec):
    grid = grid_factory(40, 30)
    rng = np.random.default_rng(5)
    ntr = 3
    qs = [rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32) for _ in range(ntr)]
    ls = [rng.standard_normal((grid.mwa, grid.mza)).astype(np.float32) for _ in range(ntr)]
    for a in qs + ls:
        a[0] = 0
    g = Context(grid, abi.make_params(prec, num_scalar=ntr, dtsm=60.0), device=0)
    try:
        for n in range(ntr):
            g.upload_scalar(n + 1, qs[n])
            g.upload_scalar_adj(n + 1, ls[n])
        for n in range(ntr):
            q = g.download_scalar(n + 1)
            lam = g.download_scalar_adj(n + 1)
            assert np.array_equal(q, qs[n]) and np.array_equal(lam, ls[n]), n
        # a tendency upload with var_p = NULL leaves the tracer alone
        g.upload_scalar(2, None, qs[0])
        assert np.array_equal(g.download_scalar(2), qs[1])
    finally:
        g.close()
