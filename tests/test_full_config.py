"""Oracle pin of the full option set together (the configuration of the reference's own
test namelist, build_olam_test/OLAMIN: 3 acoustic steps per 60 s long step, RK2
scalars, sixth-order damping, Rayleigh friction on W, divergence and top winds, a moist
atmosphere and PBL turbulence), on the CPU: through three long steps the dry mass and
every tracer's mass are conserved to rounding, which needs the density update, the
long-step mass fluxes and the tracer fluxes to stay consistent under all the options.
"""

import numpy as np

from olam_amd import abi, cases, mesh as M
from oracle import Oracle
from test_turbulence import turbulence

OLAMIN_RAYF = {"rayf": (27000.0, 0.0, 1.0), "rayfw": (27000.0, 0.5, 1.0),
               "rayfdiv": (27000.0, 0.15, 1.0), "rayfmix": (27000.0, 0.25, 1.0)}


def test_full_configuration_conserves_mass():
    grid = abi.FortranGrid(M.build_mesh(8), M.vertical_grid(30))
    st = cases.jet_state(grid)
    rng = np.random.default_rng(41)
    shape = (grid.mwa, grid.mza)
    rr_v = rng.uniform(0.0, 0.015, shape).astype(np.float32)
    rr_w = (rr_v + rng.uniform(0.0, 0.003, shape)).astype(np.float32)
    theta = (st.arrays["thil"] * (1.0 + 0.005 * rng.uniform(0.0, 1.0, shape))).astype(np.float32)
    turb = turbulence(grid, st, seed=9, drag=True)
    tracers = cases.lognormal_tracers(grid, 3)
    prm = abi.make_params("f64", num_scalar=3, nrk_wrtv=3, nrk_scal=2, order=2, dtsm=20.0, dtlm=60.0,
                          nacoust=3, divh_damp_level=3, divh_damp_fact=0.12, vort_damp_level=3,
                          vort_damp_fact=0.075, initial=2, rayf=OLAMIN_RAYF)
    o = Oracle(grid, prm)
    o.upload_state(st)
    o.init_wrtv_rk()
    for n in range(3):
        o.upload_scalar(n + 1, tracers[n], None)
    m0 = o.mass_sums()
    for _ in range(3):
        o.set_turbulence(*turb)
        o.comp_alpha_press(rr_v, rr_w, theta)
        o.tend0()
        for _ in range(prm.nacoust):
            o.prog_wrtv_rk()
        o.timeavg_momsc()
        o.scalar_transport_rk()
    m1 = o.mass_sums()
    out = abi.FortranState(grid)
    o.download_state(out)
    o.close()
    assert np.isfinite(out.arrays["wmc"]).all()
    rel = (m1[:4] - m0[:4]) / m0[:4]
    assert np.all(np.abs(rel) < 1e-13), rel
