"""Run settings: the reference's keys, defaults, derived keys and CLI.

Mirrors utils/settings.py:13-258 (load_settings): the same defaults
(:15-98), the same CLI flags (:100-107), the same dependent defaults and
validation (:132-170) and the same passive-variable offsets (:181-256).
Config files are JSON5-style (``//`` and ``/* */`` comments, trailing commas,
duplicate keys with the last one winning); the reference silently ignores the
whole file when its json5 import fails (:4-11), this loader parses it.

Additive keys (all default-off, so reference configs keep reference meaning):
  MAX_STEPS      stop after this many steps (besides T_END); 0 = unlimited
  PRECISION      "fp64" | "mixed" (FP32 reconstruction + Riemann fluxes)
  C2P_GUESS      "plain" | "learned" (RMHD: MLP initial guess for the Newton solve) |
                 "previous" (RMHD: Newton start from the stage input's primitives)
  C2P_MODEL      path of the learned-guess weights (.npz) for C2P_GUESS="learned"
  BC_Y, BC_Z     "periodic" | "outflow"
  IC_RHO_NOISE   amplitude of uniform density noise on interior cells (0 = none)
  IC_SEED        seed of that noise (numpy default_rng)
  B_AMB          ambient field [Bx, By, Bz] (interior + nozzle); None = zero
  B_PERT_AMP     rms amplitude of a divergence-free perturbation dB = curl A (0 = none)
  B_PERT_KPEAK   peak wavenumber of A's Gaussian spectrum (in units of 2 pi / L)
  B_PERT_SEED    seed of A's random modes
  NOZZLE_SIGMA   if set, B0 of a toroidal/poloidal nozzle from the magnetisation sigma
  RESISTIVE_MODE "reference" (1x MUSCL, 2x PPM, 0x WENO) | "once" (1x in every scheme)
"""
import argparse
import json
import os
import re

DEFAULTS = dict(
    # grid & numerics
    NX=128, NY=96, NZ=96, Lx=1.0, Ly=1.0, Lz=1.0,
    T_END=0.25, OUT_EVERY=50, PRINT_EVERY=10,
    NG=2, CFL=0.35, GAMMA=5.0 / 3.0,
    # jet physics
    JET_RADIUS=0.12, JET_CENTER=None,
    GAMMA_JET=6.0, ETA_RHO=0.05, P_EQ=1.0e-2,
    SHEAR_THICK=0.02,
    NOZZLE_TURB=True, TURB_VAMP=0.02, TURB_PAMP=0.00,
    NOZZLE_PROFILE="top_hat", NOZZLE_PERTURB=None,
    RECON="muscl", LIMITER="mc", RIEMANN="hlle", RK_ORDER=2,
    CHECKPOINT_EVERY=0, RESTART_PATH=None,
    # ambient
    RHO_AMB=1.0, P_AMB=None, VX_AMB=0.0, VY_AMB=0.0, VZ_AMB=0.0,
    P_MAX=1.0, V_MAX=0.999,
    # physics mode
    PHYSICS="hydro", GLM_CH=1.0, GLM_CP=0.1, B_INIT="none", B0=0.0,
    # GR placeholders
    GR_METRIC="minkowski", GR_MASS=1.0, GR_SPIN=0.0, ORTHONORMAL_FLUX=True,
    # dissipation placeholders
    DISSIPATION_ENABLED=False, DISSIPATION_MODEL="israel_stewart", RELAX_TAU=0.1,
    BULK_ZETA=0.0, SHEAR_ETA=0.0, HEAT_KAPPA=0.0,
    RELAX_TAU_BULK=None, RELAX_TAU_SHEAR=None, RELAX_TAU_HEAT=None,
    DISSIPATION_CAP_FRAC=0.5, DISSIPATION_ADVECT=True,
    # tracers
    N_TRACERS=0, TRACER_NAMES=None, TRACER_NOZZLE_VALUES=None, TRACER_AMB_VALUES=None,
    # ion chemistry
    CHEMISTRY_ENABLED=False, CHEM_X=0.76, CHEM_Y=0.24, CHEM_TUNIT_K=1.0e4, CHEM_RATE_SCALE=1.0,
    CHEM_ENERGY_SCALE=1.0, CHEM_TMIN_K=1.0, CHEM_X_HII_AMB=0.0, CHEM_X_HEII_AMB=0.0, CHEM_X_HEIII_AMB=0.0,
    CHEM_X_HII_NOZZLE=None, CHEM_X_HEII_NOZZLE=None, CHEM_X_HEIII_NOZZLE=None,
    # two-temperature
    TWO_TEMPERATURE=False, TEI_TAU=0.5, TE_AMB=None, TI_AMB=None, TE_NOZZLE=None, TI_NOZZLE=None,
    # cooling/heating
    COOLING_ENABLED=False, COOLING_LAMBDA=0.0, HEATING_RATE=0.0,
    # resistive RMHD
    RESISTIVE_ENABLED=False, RESISTIVITY=0.0,
    # results
    RESULTS_UNIQUE=False,
    # debug
    DEBUG=False, ASSERTS=False, CHECK_NAN_EVERY=0,
)

NEW_DEFAULTS = dict(
    MAX_STEPS=0, PRECISION="fp64", C2P_GUESS="plain", C2P_MODEL=None, BC_Y="periodic", BC_Z="periodic",
    IC_RHO_NOISE=0.0, IC_SEED=12345, B_AMB=None, B_PERT_AMP=0.0, B_PERT_KPEAK=8.0, B_PERT_SEED=54321,
    NOZZLE_SIGMA=None, RESISTIVE_MODE="reference",
)


def _strip_json5(text):
    """Remove // and /* */ comments (outside strings) and trailing commas."""
    out = []
    i, n = 0, len(text)
    in_str = None
    while i < n:
        ch = text[i]
        if in_str:
            out.append(ch)
            if ch == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if ch == in_str:
                in_str = None
            i += 1
            continue
        if ch in ('"', "'"):
            in_str = ch
            out.append('"')
            i += 1
            continue
        if text.startswith("//", i):
            while i < n and text[i] != "\n":
                i += 1
            continue
        if text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
            continue
        out.append(ch)
        i += 1
    s = "".join(out)
    s = re.sub(r",(\s*[}\]])", r"\1", s)
    return s


def load_config_file(path):
    """Parse a JSON5-style config file (comments, trailing commas; duplicate keys: last wins)."""
    with open(path, "r") as f:
        return json.loads(_strip_json5(f.read()))


def _expand_tracer_values(val, n, default_first, default_rest):
    # utils/settings.py:195-210
    if val is None:
        if n == 0:
            return []
        return [default_first] + [default_rest] * (n - 1)
    if isinstance(val, (int, float)):
        return [float(val)] * n
    if isinstance(val, (list, tuple)):
        out = [float(v) for v in val]
        if len(out) < n:
            out.extend([out[-1]] * (n - len(out)))
        return out[:n]
    return [default_first] + [default_rest] * max(0, n - 1)


def finalize(cfg):
    """Defaults + file values -> the reference's settings dict (utils/settings.py:123-256)."""
    s = {**DEFAULTS, **NEW_DEFAULTS, **cfg}
    if s["P_AMB"] is None:
        s["P_AMB"] = s["P_EQ"]
    if s["JET_CENTER"] is None:
        s["JET_CENTER"] = [0.0, 0.5 * s["Ly"], 0.5 * s["Lz"]]
    else:
        jc = s["JET_CENTER"]
        if isinstance(jc, dict):
            s["JET_CENTER"] = [0.0, float(jc.get("y", 0.5 * s["Ly"])), float(jc.get("z", 0.5 * s["Lz"]))]
        else:
            if not (isinstance(jc, (list, tuple)) and len(jc) == 3):
                raise ValueError("JET_CENTER must be [x,y,z] or {'y':..., 'z':...}")
            s["JET_CENTER"] = [float(jc[0]), float(jc[1]), float(jc[2])]
    if s["NG"] < 2:
        raise ValueError("NG must be >= 2 for MUSCL (uses i±2, j±2, k±2).")
    if not (0.0 < s["CFL"] <= 0.9):
        raise ValueError("CFL should be in (0, 0.9].")
    if min(s["NX"], s["NY"], s["NZ"]) < 8:
        raise ValueError("NX, NY, NZ must be ≥ 8.")
    if s.get("DISSIPATION_ENABLED", False) and s.get("PHYSICS") in ("rmhd", "grmhd"):
        raise ValueError("DISSIPATION_ENABLED is supported only for hydro/grhd.")
    if s.get("NOZZLE_PERTURB") is None:
        s["NOZZLE_PERTURB"] = "random" if s.get("NOZZLE_TURB", False) else "none"
    for key in ("NOZZLE_PROFILE", "NOZZLE_PERTURB", "LIMITER", "RECON", "RIEMANN", "PRECISION", "C2P_GUESS",
                "BC_Y", "BC_Z", "RESISTIVE_MODE"):
        if isinstance(s.get(key), str):
            s[key] = s[key].lower()
    for key in ("RELAX_TAU_BULK", "RELAX_TAU_SHEAR", "RELAX_TAU_HEAT"):
        if s.get(key) is None:
            s[key] = s.get("RELAX_TAU", 0.1)
    ntr = int(s.get("N_TRACERS", 0))
    if ntr < 0:
        raise ValueError("N_TRACERS must be >= 0.")
    s["N_TRACERS"] = ntr
    names = s.get("TRACER_NAMES")
    if isinstance(names, str):
        names = [n.strip() for n in names.split(",") if n.strip()]
    names = list(names or [])
    while len(names) < ntr:
        names.append(f"tracer{len(names)}")
    s["TRACER_NAMES"] = names[:ntr]
    s["TRACER_NOZZLE_VALUES"] = _expand_tracer_values(s.get("TRACER_NOZZLE_VALUES"), ntr, 1.0, 0.0)
    s["TRACER_AMB_VALUES"] = _expand_tracer_values(s.get("TRACER_AMB_VALUES"), ntr, 0.0, 0.0)
    if s.get("CHEMISTRY_ENABLED", False):
        s["CHEM_X"] = float(s.get("CHEM_X", 0.76))
        s["CHEM_Y"] = float(s.get("CHEM_Y", 0.24))
        for key in ("CHEM_X_HII_AMB", "CHEM_X_HEII_AMB", "CHEM_X_HEIII_AMB"):
            s[key] = max(0.0, min(1.0, float(s.get(key, 0.0))))
        for a, b in (("CHEM_X_HII_NOZZLE", "CHEM_X_HII_AMB"), ("CHEM_X_HEII_NOZZLE", "CHEM_X_HEII_AMB"),
                     ("CHEM_X_HEIII_NOZZLE", "CHEM_X_HEIII_AMB")):
            if s.get(a) is None:
                s[a] = s[b]
        s["N_CHEM"] = 3
    else:
        s["N_CHEM"] = 0
    if s.get("TWO_TEMPERATURE", False):
        tamb = s.get("P_AMB", s.get("P_EQ", 1.0e-2)) / max(s.get("RHO_AMB", 1.0), 1e-12)
        if s.get("TE_AMB") is None:
            s["TE_AMB"] = tamb
        if s.get("TI_AMB") is None:
            s["TI_AMB"] = tamb
        if s.get("TE_NOZZLE") is None:
            s["TE_NOZZLE"] = s["TE_AMB"]
        if s.get("TI_NOZZLE") is None:
            s["TI_NOZZLE"] = s["TI_AMB"]
        s["N_THERMO"] = 2
    else:
        s["N_THERMO"] = 0
    if s.get("PHYSICS") in ("rmhd", "grmhd"):
        base = 9
    else:
        base = 15 if s.get("DISSIPATION_ENABLED", False) else 5
    s["TRACER_OFFSET"] = base
    s["PASSIVE_OFFSET"] = base
    s["N_PASSIVE"] = s["N_TRACERS"] + s["N_THERMO"] + s["N_CHEM"]
    s["THERMO_OFFSET"] = base + s["N_TRACERS"]
    s["CHEM_OFFSET"] = base + s["N_TRACERS"] + s["N_THERMO"]
    s["CHEM_NAMES"] = ["xHII", "xHeII", "xHeIII"]
    return s


def build_parser():
    # utils/settings.py:100-107, plus the additive --max-steps
    ap = argparse.ArgumentParser(description="3D SRHD/RMHD jet on one MI355X (MUSCL/PPM/WENO5 + HLLE/HLLC/HLLD)")
    ap.add_argument("--config", type=str, help="path to JSON/JSON5 config")
    ap.add_argument("--debug", action="store_true", help="verbose logging")
    ap.add_argument("--nx", type=int)
    ap.add_argument("--ny", type=int)
    ap.add_argument("--nz", type=int)
    ap.add_argument("--t-end", type=float)
    ap.add_argument("--out-every", type=int)
    ap.add_argument("--print-every", type=int)
    ap.add_argument("--max-steps", type=int)
    return ap


def load_settings(argv=None):
    """Parse the CLI, read --config, merge, derive.  Same keys as the reference."""
    args = build_parser().parse_args(argv)
    cfg = {}
    if args.config:
        if not os.path.exists(args.config):
            raise FileNotFoundError(args.config)
        cfg = load_config_file(args.config)
    s = {**DEFAULTS, **NEW_DEFAULTS, **cfg}
    if args.debug:
        s["DEBUG"] = True
    for attr, key in (("nx", "NX"), ("ny", "NY"), ("nz", "NZ"), ("t_end", "T_END"), ("out_every", "OUT_EVERY"),
                      ("print_every", "PRINT_EVERY"), ("max_steps", "MAX_STEPS")):
        v = getattr(args, attr)
        if v is not None:
            s[key] = v
    return finalize(s)
