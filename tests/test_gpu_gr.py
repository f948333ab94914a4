"""GPU tests of the GRHD / GRMHD wrappers (core/grhd_core.py, core/grmhd_core.py, core/gr_metric.py)
against the reference's outputs (tests/golden/gr.npz) and the CPU oracle.

The wrappers run the flat RHS on a copy of the state whose interior velocities are divided by the
Kerr-Schild lapse (ORTHONORMAL_FLUX with a Schwarzschild or Kerr-Schild metric) and add the
lapse-gradient sources from the unscaled state (every metric but Minkowski).  Bars: the RHS
<= 1e-11 relative and steps <= 1e-10 relative per variable, those of the flat RHS and step tests
(the flux kernels' FMA contraction; the sources themselves are evaluated without contraction)."""
import ast
import os

import numpy as np
import pytest

import oracle as O
from tests.gpu_helpers import log, make_domain, rel_err

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _domain(case, pr, ng, dxyz, rk=2):
    phys, metric, ortho, mass, gamma, recon, lim, riem = case
    rm = phys == "grmhd"
    return make_domain("rmhd" if rm else "hydro", pr.shape, ng, dxyz, gamma=gamma if rm else None, limiter=lim,
                       recon=recon, riemann=riem, rk_order=rk, gr_metric=metric, gr_mass=mass,
                       orthonormal_flux=ortho, gr_gamma=gamma)


def test_gr_rhs_and_steps_golden():
    f = np.load(os.path.join(G, "gr.npz"))
    dxyz = tuple(float(v) for v in f["dxyz"])
    ng = int(f["ng"])
    cases = [ast.literal_eval(str(c)) for c in f["rhs_cases"]]
    for c, case in enumerate(cases):
        pr = f[f"rhs_in_{c}"]
        log(f"[gpu] gr rhs {case}")
        with _domain(case, pr, ng, dxyz) as d:
            d.set_prim(pr)
            rhs = d.compute_rhs()
        err = rel_err(rhs, f[f"rhs_{c}"])
        log(f"      rel err {err:.3e}")
        assert err <= 1e-11, case
    for c, (ci, rk) in enumerate(f["step_cases"]):
        case = cases[ci]
        pr = f[f"rhs_in_{ci}"]
        log(f"[gpu] gr step {case} rk{rk}")
        with _domain(case, pr, ng, dxyz, rk=int(rk)) as d:
            d.set_prim(pr)
            d.step(float(f["dt"]))
            out = d.get_prim()
        err = max(rel_err(out[v], f[f"step_{c}"][v]) for v in range(out.shape[0]))
        log(f"      rel err {err:.3e}")
        assert err <= 1e-10, case


def test_gr_compat_modules_vs_oracle():
    """grhd_core / grmhd_core (reference names) on larger states against the oracle."""
    from astrofd_amd.core import grhd_core, grmhd_core, rmhd_core, srhd_core
    from tests.golden_states import rmhd_state, srhd_state
    shape = (20, 18, 16)
    dx, dy, dz = 1 / 14, 1 / 12, 1 / 10
    base = dict(NG=3, P_MAX=2.0, V_MAX=0.9999, LIMITER="mc", N_PASSIVE=0)
    srhd_core.configure(dict(base, PHYSICS="hydro", GAMMA=1.4, RECON="ppm", RIEMANN="hllc", PASSIVE_OFFSET=5))
    grhd_core.configure(dict(GR_METRIC="schwarzschild", GR_MASS=0.08, ORTHONORMAL_FLUX=True))
    pr = srhd_state(shape, 71, jet=False)
    log("[gpu] grhd compat vs oracle")
    got = grhd_core.compute_rhs_grhd(pr, dx, dy, dz, 0, 3)
    P = O.make_params("hydro", shape, ng=3, p_max=2.0, v_max=0.9999, recon="ppm", riemann="hllc",
                      gr_metric="schwarzschild", gr_mass=0.08, gr_gamma=1.4)
    ref = O.rhs(P, pr, dx, dy, dz)
    assert rel_err(got, ref) <= 1e-11
    out = grhd_core.step_ssprk3(pr, dx, dy, dz, 0.01, 0, 3)
    P.rk_order = 3
    assert max(rel_err(out[v], r) for v, r in enumerate(O.step(P, pr, dx, dy, dz, 0.01))) <= 1e-10
    rmhd_core.configure(dict(base, PHYSICS="rmhd", GAMMA=4.0 / 3.0, RECON="muscl", RIEMANN="hlld", PASSIVE_OFFSET=9,
                             GLM_CH=1.0, GLM_CP=0.1))
    grmhd_core.configure(dict(GR_METRIC="kerr-schild", GR_MASS=0.08, ORTHONORMAL_FLUX=True))
    pr = rmhd_state(shape, 72, jet=False)
    log("[gpu] grmhd compat vs oracle")
    got = grmhd_core.compute_rhs_grmhd(pr, dx, dy, dz, 0, 3)
    P = O.make_params("grmhd", shape, ng=3, gamma=4.0 / 3.0, p_max=2.0, v_max=0.9999, recon="muscl",
                      riemann="hlld", gr_metric="kerr-schild", gr_mass=0.08)
    assert rel_err(got, O.rhs(P, pr, dx, dy, dz)) <= 1e-11
