"""GPU tests of the post-step sources: cooling/heating (core/source_terms.py:5-28), ion chemistry
(core/chemistry.py:43-139), the fused pass of the driver (cooling, two-temperature, chemistry,
solvers/srhd3d_mpi_muscl.py:324-326) and the nozzle's chemistry fill (core/nozzle.py:77-82),
against the reference's outputs in tests/golden/sources.npz and the CPU oracle.

Bars: cooling, two-temperature and the nozzle are +, -, *, / and compares, compiled without FMA
contraction: bitwise.  Chemistry evaluates exp, pow and sqrt; the device's exp and pow may differ
from glibc's in the last bit (both are faithfully rounded, <= 1 ulp), which the IMEX update
carries to <= ~10 ulp of the fractions and of p when dp does not cancel p; the bar is 1e-13
relative to max(|result|, |p_in|) for p and 1e-13 absolute for the fractions (which lie in [0, 1]),
and the floors (p = 1e-20, clamps to 0 and 1) must agree exactly."""
import numpy as np
import pytest

import oracle as O
from tests import sources_cases as S
from tests.gpu_helpers import log, make_domain

pytestmark = pytest.mark.gpu


def _domain(pr, ng, two_temp=None):
    nvb = 5
    kw = dict(n_passive=pr.shape[0] - nvb)
    if two_temp is not None:
        kw.update(two_temperature=True, thermo_offset=two_temp[0], tei_tau=two_temp[1])
    return make_domain("hydro", pr.shape, ng, **kw)


def _chem_close(got, want, p_in, what):
    """got/want: [p, xHII, xHeII, xHeIII] over the grid."""
    scale = np.maximum(np.abs(want[0]), np.abs(p_in))
    ep = float((np.abs(got[0] - want[0]) / np.where(scale > 0, scale, 1.0)).max())
    ex = float(np.abs(got[1:] - want[1:]).max())
    floors = (want[0] == 1e-20) | (got[0] == 1e-20)
    log(f"      {what}: p rel err {ep:.2e}  fractions abs err {ex:.2e}  floors {int(floors.sum())}")
    assert ep <= 1e-13 and ex <= 1e-13, what
    assert np.array_equal(got[0][floors], want[0][floors]), what
    for m in (1, 2, 3):
        for edge in (0.0, 1.0):
            assert np.array_equal(got[m] == edge, want[m] == edge), f"{what}: clamp at {edge}, fraction {m}"


def test_cooling_heating_golden():
    f = S.load()
    for c, dt, cfg, pr, out in S.cooling_cases(f):
        log(f"test_cooling_heating_golden case {c} dt={dt} {cfg}")
        with _domain(pr, cfg["NG"]) as d:
            d.set_prim(pr)
            if cfg.get("COOLING_ENABLED", False):
                d.cooling_heating(dt, cfg.get("COOLING_LAMBDA", 0.0), cfg.get("HEATING_RATE", 0.0),
                                  cfg.get("P_MAX", 1.0))
            got = d.get_prim()
        assert np.array_equal(got[4], out), f"cooling {c}"
        assert np.array_equal(got[:4], pr[:4])


def test_ion_chemistry_golden():
    f = S.load()
    for c, dt, cfg, pr, out, nvb in S.chem_cases(f):
        log(f"test_ion_chemistry_golden case {c} dt={dt} nv={pr.shape[0]}")
        with _domain(pr, cfg["NG"]) as d:
            d.set_prim(pr)
            d.ion_chemistry(dt, cfg)
            got = d.get_prim()
        _chem_close(got[[4, nvb, nvb + 1, nvb + 2]], out, pr[4], f"chemistry {c}")
        untouched = [v for v in range(pr.shape[0]) if v not in (4, nvb, nvb + 1, nvb + 2)]
        assert np.array_equal(got[untouched], pr[untouched])


def test_post_step_sources_chain_golden():
    """The fused pass (one kernel, cooling -> two-temperature -> chemistry per cell) against the
    reference's three calls in sequence."""
    from astrofd_amd.domain import sources_from_settings
    f = S.load()
    pr = f["chain_in"]
    cfg = dict(S.CHAIN_CFG, NG=int(f["ng"]), N_PASSIVE=6)
    log("test_post_step_sources_chain_golden")
    with _domain(pr, cfg["NG"], two_temp=(cfg["THERMO_OFFSET"], cfg["TEI_TAU"])) as d:
        d.set_prim(pr)
        d.post_step_sources(S.CHAIN_DT, sources_from_settings(cfg))
        got = d.get_prim()
    want = f["chain_out"]  # p, Te, Ti, xHII, xHeII, xHeIII
    assert np.array_equal(got[6:8], want[1:3])  # Te, Ti read p after cooling only: bitwise
    _chem_close(got[[4, 8, 9, 10]], want[[0, 3, 4, 5]], pr[4], "chain")


def test_nozzle_chemistry_golden():
    f = S.load()
    for c, name, cfg, pr, out, (NX, NY, NZ, ng) in S.nozzle_cases(f):
        log(f"test_nozzle_chemistry_golden {name}")
        s = dict(cfg, NX=NX, NY=NY, NZ=NZ, Lx=1.0, Ly=1.0, Lz=1.0, JET_CENTER=[0.0, 0.45, 0.55])
        phys = cfg["PHYSICS"]
        nvb = 9 if phys == "rmhd" else 5
        with make_domain(phys, pr.shape, ng, (1.0 / NX, 1.0 / NY, 1.0 / NZ), n_passive=pr.shape[0] - nvb,
                         two_temperature=cfg["TWO_TEMPERATURE"], thermo_offset=cfg["THERMO_OFFSET"]) as d:
            d.set_prim(pr)
            d.set_nozzle(s)
            d.apply_nozzle_left_x(np.random.default_rng(1234))
            got = d.get_prim()
        assert np.array_equal(got[:, :ng], out), name
        assert np.array_equal(got[:, ng:], pr[:, ng:]), name


def test_sources_compat_modules_vs_oracle():
    """The reference-named modules (in place on numpy arrays) against the oracle on a larger state."""
    from astrofd_amd.core import chemistry, source_terms
    from tests.golden_states import source_state
    ng = 3
    shape = (20, 18, 16)
    pr = source_state(shape, 9, 11, chem_off=8)
    cfg = dict(S.CHAIN_CFG, NG=ng)
    log("test_sources_compat_modules_vs_oracle")
    got = pr.copy()
    source_terms.apply_cooling_heating(got, 0.03, cfg)
    source_terms.apply_two_temperature(got, 0.03, cfg)
    chemistry.apply_ion_chemistry(got, 0.03, cfg)
    ref = pr.copy()
    P = O.make_params("hydro", shape, nv=11, ng=ng)
    O.apply_cooling_heating(P, ref, 0.03, cfg)
    O.apply_two_temperature(P, ref, 0.03, cfg["TEI_TAU"], cfg["THERMO_OFFSET"])
    O.apply_ion_chemistry(P, ref, 0.03, cfg)
    assert np.array_equal(got[[0, 1, 2, 3, 5, 6, 7]], ref[[0, 1, 2, 3, 5, 6, 7]])
    _chem_close(got[[4, 8, 9, 10]], ref[[4, 8, 9, 10]], pr[4], "compat")
