"""GPU tests of Israel-Stewart dissipation against the reference's outputs
(tests/golden/dissipation.npz, driver_diss_10.npz) and the CPU oracle:

- the SRHD RHS with the dissipative hooks of core/srhd_core.py:393-462 (bulk pressure on the
  MUSCL face pressures; Pi, pi^ij, q^i fluxes once per MUSCL face, twice per PPM face, none in
  WENO), PASSIVE_OFFSET = 15 with two passives: <= 1e-11 relative, the bar of the other RHS tests;
- the step carrying fields 5..14 unchanged through the stages: <= 1e-10 relative per variable;
- core/dissipation.py's post-step operator (compiled without FMA contraction, divisions and
  compares only): bitwise;
- the reference-named module and the fused post-step entry point against the oracle.
"""
import ast
import os

import numpy as np
import pytest

import oracle as O
from tests.gpu_helpers import log, make_domain, rel_err

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    return np.load(os.path.join(G, "dissipation.npz"))


def per_var_rel(a, b):
    return max(rel_err(a[v], b[v]) for v in range(a.shape[0]))


def test_dissipative_rhs_golden():
    f = load()
    dx, dy, dz = f["dxyz"]
    pr = f["rhs_in"]
    for c, name in enumerate(f["rhs_cases"]):
        recon, lim, riem = str(name).split("/")
        log(f"[gpu] dissipative rhs {name}")
        with make_domain("hydro", pr.shape, 3, (dx, dy, dz), limiter=lim, recon=recon, riemann=riem, n_passive=2,
                         passive_offset=15) as d:
            d.set_prim(pr)
            rhs = d.compute_rhs()
        err = rel_err(rhs, f[f"rhs_{c}"])
        log(f"      rel err {err:.3e}")
        assert err <= 1e-11, name


def test_dissipative_step_golden():
    f = load()
    dx, dy, dz = f["dxyz"]
    pr = f["step_in"]
    for c, name in enumerate(f["step_cases"]):
        recon, lim, riem, rk = str(name).split("/")
        log(f"[gpu] dissipative step {name}")
        with make_domain("hydro", pr.shape, 3, (dx, dy, dz), limiter=lim, recon=recon, riemann=riem, n_passive=2,
                         passive_offset=15, rk_order=int(rk)) as d:
            d.set_prim(pr)
            d.step(float(f["dt"]))
            out = d.get_prim()
        err = per_var_rel(out, f[f"step_{c}"])
        log(f"      rel err {err:.3e}")
        assert err <= 1e-10, name
        assert np.array_equal(out[5:15], pr[5:15])  # carried through the stages


def test_dissipation_operator_golden():
    from astrofd_amd.domain import diss_from_settings
    f = load()
    dx, dy, dz = f["dxyz"]
    pr = f["op_in"]
    for c, s in enumerate(f["op_cases"]):
        enabled, cfg = ast.literal_eval(str(s))
        log(f"[gpu] dissipation operator case {c} enabled={enabled} {cfg}")
        with make_domain("hydro", pr.shape, 3, (dx, dy, dz), n_passive=pr.shape[0] - 5) as d:
            d.set_prim(pr)
            if enabled:
                d.causal_dissipation(0.01, diss_from_settings(cfg))
            out = d.get_prim()
        assert np.array_equal(out[5:15], f[f"op_{c}"]), f"operator case {c}"
        assert np.array_equal(out[:5], pr[:5])


def test_dissipation_compat_and_fused_vs_oracle():
    """core.dissipation.apply_causal_dissipation (in place on numpy) and afd_post_step_sources
    (dissipation, then cooling) against the oracle on a larger state with a W ~ 7 jet."""
    from astrofd_amd.core import dissipation
    from astrofd_amd.domain import sources_from_settings
    from tests.golden_states import srhd_state
    shape = (22, 20, 18)
    base = srhd_state(shape, 61, npass=1)
    diss = np.random.default_rng(62).uniform(-0.004, 0.004, (10,) + shape)
    pr = np.concatenate([base[:5], diss, base[5:]])
    dx, dy, dz = 1 / 16, 1 / 14, 1 / 12
    cfg = dict(DISSIPATION_ENABLED=True, PHYSICS="hydro", NG=3, BULK_ZETA=0.02, SHEAR_ETA=0.01, HEAT_KAPPA=0.005,
               RELAX_TAU=0.05, P_MAX=2.0, COOLING_ENABLED=True, COOLING_LAMBDA=0.5, HEATING_RATE=0.001)
    log("[gpu] dissipation compat + fused vs oracle")
    got = dissipation.apply_causal_dissipation(pr.copy(), 0.004, dx, dy, dz, cfg)
    ref = pr.copy()
    P = O.make_params("hydro", shape, nv=16, ng=3, n_passive=1, passive_offset=15)
    O.apply_causal_dissipation(P, ref, 0.004, dx, dy, dz, cfg)
    assert np.array_equal(got, ref)
    with make_domain("hydro", pr.shape, 3, (dx, dy, dz), n_passive=1, passive_offset=15) as d:
        d.set_prim(pr)
        d.post_step_sources(0.004, sources_from_settings(cfg))
        got = d.get_prim()
    O.apply_cooling_heating(P, ref, 0.004, cfg)
    assert np.array_equal(got, ref)
