"""CPU tests of the host side: settings (utils/settings.py), initial state (the fixture's own
init), the divergence-free field perturbation, the nozzle profile, driver helpers, and the
reference-named modules' argument handling (no GPU calls)."""
import ast
import os

import numpy as np
import pytest

import oracle
from astrofd_amd import ic
from astrofd_amd.core import _ctx
from astrofd_amd.core import nozzle as cnozzle
from astrofd_amd.settings import DEFAULTS, _strip_json5, finalize, load_config_file, load_settings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
CFG = os.path.join(ROOT, "astrofd_amd", "config")


def test_defaults_and_derived_keys():
    s = finalize({})
    assert s["NX"] == DEFAULTS["NX"] and s["P_AMB"] == s["P_EQ"]
    assert s["JET_CENTER"] == [0.0, 0.5, 0.5]
    assert s["NOZZLE_PERTURB"] == "random"  # NOZZLE_TURB defaults to True
    assert (s["TRACER_OFFSET"], s["PASSIVE_OFFSET"], s["N_PASSIVE"]) == (5, 5, 0)
    r = finalize(dict(PHYSICS="rmhd", N_TRACERS=2, TWO_TEMPERATURE=True))
    assert (r["TRACER_OFFSET"], r["THERMO_OFFSET"], r["N_PASSIVE"]) == (9, 11, 4)
    assert r["TRACER_NAMES"] == ["tracer0", "tracer1"] and r["TRACER_NOZZLE_VALUES"][0] == 1.0
    assert r["TE_AMB"] == pytest.approx(0.01)
    d = finalize(dict(DISSIPATION_ENABLED=True))
    assert d["PASSIVE_OFFSET"] == 15
    with pytest.raises(ValueError):
        finalize(dict(NG=1))
    with pytest.raises(ValueError):
        finalize(dict(CFL=1.5))
    with pytest.raises(ValueError):
        finalize(dict(JET_CENTER=[0.5, 0.5]))


def test_json5_and_cli(tmp_path):
    text = '{\n  // comment\n  "NX": 16, /* block */ "NY": 12,\n  "NZ": 10, "V_MAX": 0.9, "V_MAX": 0.99,\n}\n'
    assert _strip_json5('{"a": "x // not a comment"}') == '{"a": "x // not a comment"}'
    p = tmp_path / "c.json"
    p.write_text(text)
    cfg = load_config_file(str(p))
    assert cfg == {"NX": 16, "NY": 12, "NZ": 10, "V_MAX": 0.99}  # last duplicate wins
    s = load_settings(["--config", str(p), "--nx", "20", "--t-end", "0.5", "--max-steps", "3", "--debug"])
    assert (s["NX"], s["NY"], s["T_END"], s["MAX_STEPS"], s["DEBUG"]) == (20, 12, 0.5, 3, True)


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(CFG) if f.endswith(".json")))
def test_shipped_configs_parse(name):
    s = finalize(load_config_file(os.path.join(CFG, name)))
    assert s["PHYSICS"] in ("hydro", "rmhd")


def test_initial_state_matches_fixture():
    """The 32^3 SRHD jet's initial state: bitwise the fixture's (reference init_block + noise)."""
    f = np.load(os.path.join(G, "driver_srhd_32.npz"))
    s = finalize(load_config_file(os.path.join(CFG, "srhd_jet_32.json")))
    pr = ic.initial_state(s)
    assert np.array_equal(pr, f["pr0"])


def test_field_perturbation_is_divergence_free():
    bx, by, bz = ic.vector_potential_field((24, 20, 16), 4.0, 7, 1e-3)
    rms = np.sqrt(np.mean(bx ** 2 + by ** 2 + bz ** 2))
    assert rms == pytest.approx(1e-3, rel=1e-12)

    def d(f, axis, n):
        return (np.roll(f, -1, axis) - np.roll(f, 1, axis)) * n / 2.0

    div = d(bx, 0, 24) + d(by, 1, 20) + d(bz, 2, 16)
    assert np.abs(div).max() < 1e-14 * np.abs(bx).max() * 24
    s = finalize(dict(PHYSICS="rmhd", NX=16, NY=16, NZ=16, NG=3, GAMMA=4 / 3, B_AMB=[0.1, 0, 0], B_PERT_AMP=1e-3,
                      NOZZLE_SIGMA=0.1, GAMMA_JET=7.0888121, ETA_RHO=0.01, P_EQ=0.01))
    pr = ic.initial_state(s)
    assert s["B0"] == pytest.approx(7.0888121 * np.sqrt(0.1 * 0.01 * 5.0))
    assert abs(np.mean(pr[5, 3:-3, 3:-3, 3:-3]) - 0.1) < 1e-4


@pytest.mark.parametrize("profile", ["top_hat", "taper", "parabolic"])
@pytest.mark.parametrize("shear", [0.0, 0.02])
def test_nozzle_weight_vs_oracle(profile, shear):
    for r in np.linspace(0.0, 0.3, 61):
        assert cnozzle.nozzle_weight(r, 0.125, shear, profile) == oracle.nozzle_weight(r, 0.125, shear, profile)


def test_driver_restart_helpers(tmp_path):
    from astrofd_amd import driver
    pr = np.random.default_rng(0).random((5, 6, 6, 6))
    ck = tmp_path / "checkpoint_rank0000_step000010.npz"
    np.savez(ck, prim=pr, meta=np.array([0.1, 0.1, 0.1, 0.25, 10.0]), comment="Checkpoint (with ghosts).")
    got, t, step = driver.load_checkpoint(str(ck))
    assert np.array_equal(got, pr) and t == 0.25 and step == 10
    assert driver.resolve_restart_path(str(tmp_path)) == str(ck)
    assert driver.resolve_restart_path(str(ck)) == str(ck)
    assert driver.resolve_restart_path("none") is None and driver.resolve_restart_path(None) is None
    s = finalize(dict(PHYSICS="rmhd", N_TRACERS=1, TWO_TEMPERATURE=True))
    f = driver.output_fields(s, np.zeros((12, 4, 4, 4)))
    assert set(f) == {"rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "psi", "tracer0", "tracer_names", "Te", "Ti"}


def test_compat_requires_configure():
    _ctx._SETTINGS["hydro"] = None
    with pytest.raises(RuntimeError):
        _ctx.settings("hydro")
    q, shape = _ctx.as_batch((1.0, 0.1, 0.0, 0.0, 0.01), 5)
    assert q.shape == (1, 5) and shape == ()
    q, shape = _ctx.as_batch((np.ones((2, 3)), 0.1, 0.0, 0.0, 0.01), 5)
    assert q.shape == (6, 5) and shape == (2, 3)
    out = _ctx.unbatch(q, shape)
    assert len(out) == 5 and out[0].shape == (2, 3)


def test_fixture_settings_readable():
    f = np.load(os.path.join(G, "driver_rmhd_8.npz"))
    s = ast.literal_eval(str(f["settings"]))
    assert s["PHYSICS"] == "rmhd" and s["RIEMANN"] == "hlld_full"


def test_diagnostics_oracle_closed_forms():
    """The diagnostics oracle on fields with known answers: a uniform moving state (max Gamma and
    inlet flux in closed form) and a linear B field (div B = sum of its slopes everywhere)."""
    from oracle import diagnostics as D
    ng, n = 3, (10, 8, 6)
    pr = np.zeros((9,) + tuple(m + 2 * ng for m in n))
    pr[0], pr[1], pr[4] = 2.0, 0.6, 0.1
    x = np.arange(pr.shape[1])[:, None, None] * 0.1
    y = np.arange(pr.shape[2])[None, :, None] * 0.1
    z = np.arange(pr.shape[3])[None, None, :] * 0.1
    pr[5] = 3.0 * x + 0 * y + 0 * z
    pr[6] = -1.0 * y + 0 * x + 0 * z
    pr[7] = 0.5 * z + 0 * x + 0 * y
    d = D.diagnostics(pr, ng, 0.1, 0.1, 0.1, rmhd=True, p_max=2.0, v_max=0.9999)
    W = 1.0 / np.sqrt(1.0 - 0.36)
    assert abs(d["max_gamma"] - W) < 1e-14
    assert abs(d["divb_mean"] - 2.5) < 1e-12 and abs(d["divb_max"] - 2.5) < 1e-12
    assert abs(d["divb_rms"] - 2.5) < 1e-12
    # SRHD inlet flux: Sx = rho h W^2 vx (Gamma 5/3), summed over the ny*nz inlet cells
    h = 1.0 + 2.5 * 0.1 / 2.0
    ds = D.diagnostics(pr[:5], ng, 0.1, 0.1, 0.1, rmhd=False)
    assert abs(ds["inlet_flux"] - 2.0 * h * W * W * 0.6 * 0.01 * n[1] * n[2]) < 1e-12
    c = D.centerline(pr, ng, 0.9999)
    assert c.shape == (4, n[0]) and np.allclose(c[0], W) and np.allclose(c[3], 0.6)


def test_post_step_source_settings():
    """The settings gates of the post-step sources (core/source_terms.py:6-11,
    core/chemistry.py:48-57) and the chemistry layout (solvers/srhd3d_mpi_muscl.py:83-84, 159-162)."""
    from astrofd_amd.domain import nozzle_from_settings, sources_from_settings
    s = finalize(dict(DEFAULTS, NX=8, NY=8, NZ=8, N_TRACERS=1, TWO_TEMPERATURE=True, CHEMISTRY_ENABLED=True,
                      CHEM_X_HII_AMB=1.7, CHEM_X_HEII_AMB=0.2, COOLING_ENABLED=True, COOLING_LAMBDA=0.3,
                      GAMMA=1.4))
    assert s["N_CHEM"] == 3 and s["CHEM_OFFSET"] == 5 + 1 + 2 and s["N_PASSIVE"] == 6
    assert s["CHEM_X_HII_AMB"] == 1.0 and s["CHEM_X_HII_NOZZLE"] == 1.0  # clamped, nozzle defaults to ambient
    src = sources_from_settings(s)
    assert src.cooling == 1 and src.cool == 0.3 and src.heat == 0.0 and src.p_cap == s["P_MAX"]
    assert src.chemistry == 1 and src.chem.offset == 8 and src.chem.gamma == 1.4 and src.chem.tunit_k == 1e4
    n = nozzle_from_settings(s)
    assert n.chemistry == 1 and n.chem_offset == 8 and list(n.chem_nozzle) == [1.0, 0.2, 0.0]
    pr = ic.init_block(s)
    assert pr.shape[0] == 11 and (pr[8] == 1.0).all() and (pr[9] == 0.2).all() and (pr[10] == 0.0).all()
    off = sources_from_settings(dict(s, COOLING_ENABLED=False, N_CHEM=2))
    assert off.cooling == 0 and off.chemistry == 0


def test_params_for_gr_and_dissipation():
    """PHYSICS grhd / grmhd map to the SRHD / RMHD context with the GR fields set (core/grhd_core.py:13-18);
    DISSIPATION_ENABLED moves the passives to PASSIVE_OFFSET = 15 (utils/settings.py:250-252)."""
    from astrofd_amd.domain import gr_metric_id, params_from_settings, sources_from_settings
    s = finalize(dict(DEFAULTS, NX=8, NY=8, NZ=8, PHYSICS="grhd", GAMMA=1.4, GR_METRIC="Schwarzschild",
                      GR_MASS=0.3, ORTHONORMAL_FLUX=False))
    p = params_from_settings(s)
    assert p.physics == 0 and p.gr_metric == 1 and p.gr_orthonormal == 0 and p.gr_mass == 0.3
    assert p.gamma == 5.0 / 3.0 and p.gr_gamma == 1.4  # Q5: the SRHD core keeps 5/3, the GR source takes GAMMA
    s = finalize(dict(DEFAULTS, NX=8, NY=8, NZ=8, PHYSICS="grmhd", GAMMA=4.0 / 3.0, GR_METRIC="kerr-schild"))
    p = params_from_settings(s)
    assert p.physics == 1 and p.gr_metric == 2 and p.gr_gamma == p.gamma == 4.0 / 3.0
    assert gr_metric_id("minkowski") == 0 and gr_metric_id("kerr") == 3
    s = finalize(dict(DEFAULTS, NX=8, NY=8, NZ=8, PHYSICS="hydro", DISSIPATION_ENABLED=True, N_TRACERS=1,
                      BULK_ZETA=0.1, RELAX_TAU=0.2, RELAX_TAU_SHEAR=0.0))
    p = params_from_settings(s)
    assert p.passive_offset == 15 and p.n_passive == 1 and p.nv == 16
    src = sources_from_settings(s)
    assert src.dissipation == 1 and src.diss.zeta == 0.1 and src.diss.tau_bulk == 0.2 and src.diss.tau_shear == 0.2
    assert ic.init_block(s).shape[0] == 16
    with pytest.raises(NotImplementedError):
        params_from_settings(dict(s, PHYSICS="mhd"))


def test_multiply_shift_division_is_exact():
    """The kernels' index division (UDiv in csrc/afd_kernels.hpp): with s = ceil(log2 d) and
    m = floor(2^(31+s) / d) + 1, (n m) >> (31 + s) equals n / d for every 0 <= n < 2^31, and m fits
    in 32 bits.  Checked by brute force where an error would first show: at both sides of every
    multiple of d near 0 and near 2^31, for all d <= 4096 and for sampled d up to 2^31 - 1."""
    rng = np.random.default_rng(5)
    ds = list(range(1, 4097)) + [int(x) for x in rng.integers(4097, 2**31 - 1, 400)] + [2**31 - 1, 2**30, 2**30 + 1]
    top = np.int64(2**31 - 1)
    for d in ds:
        s = 0
        while (1 << s) < d:
            s += 1
        m = (1 << (31 + s)) // d + 1
        assert m < 2**32
        qs = np.unique(np.concatenate([np.arange(0, 64), top // d - np.arange(0, 64)]))
        qs = qs[qs >= 0].astype(object)
        n = np.concatenate([qs * d, qs * d - 1, qs * d + d - 1])
        n = np.array([int(x) for x in n if 0 <= x <= int(top)], dtype=object)
        got = (n * m) >> (31 + s)
        assert np.array_equal(got, n // d), d
