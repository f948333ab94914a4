"""The post-step source cases of tests/golden/sources.npz, unpacked for the oracle and GPU tests.

Fixture layout (tests/golden/make_golden.py, gen_sources):
- cooling (core/source_terms.py:5-28): full input, output p;
- ion chemistry (core/chemistry.py:43-139): full input, output p and the three fractions;
- chain (cooling, two-temperature, chemistry, the order of solvers/srhd3d_mpi_muscl.py:324-326):
  full input, output p, Te, Ti and the fractions;
- nozzle with chemistry fractions (core/nozzle.py:27-92): full input, output -x ghost planes.
"""
import ast
import os

import numpy as np

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CHAIN_CFG = dict(COOLING_ENABLED=True, COOLING_LAMBDA=2.0, HEATING_RATE=0.05, P_MAX=50.0,
                 TWO_TEMPERATURE=True, N_TRACERS=1, THERMO_OFFSET=6, TEI_TAU=0.2,
                 CHEMISTRY_ENABLED=True, N_CHEM=3, CHEM_OFFSET=8, CHEM_TUNIT_K=1.0e6, CHEM_RATE_SCALE=1.0e9,
                 CHEM_ENERGY_SCALE=1.0e-8)
CHAIN_DT = 0.02


def load():
    return np.load(os.path.join(G, "sources.npz"))


def cooling_cases(f):
    """(index, dt, cfg, pr_in, p_out)"""
    for c, s in enumerate(f["cool_cases"]):
        dt, cfg = ast.literal_eval(str(s))
        yield c, dt, dict(cfg, NG=int(f["ng"])), f[f"cool_in_{c}"], f[f"cool_out_{c}"]


def chem_cases(f):
    """(index, dt, cfg, pr_in, [p, xHII, xHeII, xHeIII] out, nvb)"""
    for c, s in enumerate(f["chem_cases"]):
        dt, cfg = ast.literal_eval(str(s))
        nvb = 9 if c % 2 else 5
        cfg = dict(cfg, CHEMISTRY_ENABLED=True, N_CHEM=3, CHEM_OFFSET=nvb, NG=int(f["ng"]))
        yield c, dt, cfg, f[f"chem_in_{c}"], f[f"chem_out_{c}"], nvb


def nozzle_cases(f):
    """(index, name, cfg, pr_in, ghost-plane out, geometry (NX, NY, NZ, ng))"""
    ng = int(f["ng"])
    NX, NY, NZ = 2, 6, 5
    for c, name in enumerate(f["noz_cases"]):
        phys, ntr, tt, coff, nv, binit = str(name).split("/")
        ntr, coff, nv, tt = int(ntr), int(coff), int(nv), tt == "True"
        nvb = 9 if phys == "rmhd" else 5
        cfg = dict(NG=ng, JET_RADIUS=0.3, SHEAR_THICK=0.02, NOZZLE_PROFILE="top_hat", GAMMA_JET=7.0888121,
                   ETA_RHO=0.01, P_EQ=0.01, RHO_AMB=1.0, VX_AMB=0.01, VY_AMB=-0.02, VZ_AMB=0.03,
                   NOZZLE_PERTURB="random", TURB_VAMP=0.02, TURB_PAMP=0.1, N_TRACERS=ntr, TRACER_OFFSET=nvb,
                   TRACER_NOZZLE_VALUES=[1.0, 0.5, 0.2][:ntr], TRACER_AMB_VALUES=[0.0, 0.25, 0.1][:ntr],
                   TWO_TEMPERATURE=tt, THERMO_OFFSET=nvb + ntr, TE_NOZZLE=0.7, TI_NOZZLE=0.9,
                   CHEMISTRY_ENABLED=True, CHEM_OFFSET=coff, CHEM_X_HII_NOZZLE=0.9, CHEM_X_HEII_NOZZLE=0.3,
                   CHEM_X_HEIII_NOZZLE=0.05, PHYSICS=phys, B_INIT=binit, B0=0.5)
        pr = f[f"noz_in_{c}"]
        assert pr.shape[0] == nv
        yield c, str(name), cfg, pr, f[f"noz_out_{c}"], (NX, NY, NZ, ng)
