"""Initial state of a run: the reference's init_block plus the additive benchmark keys.

init_block restates solvers/srhd3d_mpi_muscl.py:128-163 (ambient fill of every
cell, ghosts included).  The additive keys (default off) then modify it:
  IC_RHO_NOISE / IC_SEED  rho *= 1 + U(-a, a) on interior cells, drawn with
                          numpy default_rng(IC_SEED).uniform over (NX, NY, NZ), C order
  B_AMB                   uniform field added on every cell (RMHD)
  B_PERT_*                dB = curl A on interior cells; A has random complex Gaussian
                          Fourier modes weighted by exp(-(|k| - k_peak)^2 / 8) (|k| in units
                          of 2 pi / L, k = 0 removed); the curl is the periodic 2nd-order
                          centred one, so the centred divergence of dB vanishes to round-off;
                          scaled to rms |dB| = B_PERT_AMP
  NOZZLE_SIGMA            B0 of the nozzle field from the magnetisation sigma = b^2 / (rho h)
                          of the jet: B0 = W_jet sqrt(sigma rho_jet h_jet)  (B perpendicular to v)
"""
import math

import numpy as np


def nozzle_b0_from_sigma(s):
    sigma = float(s["NOZZLE_SIGMA"])
    g = float(s.get("GAMMA", 5.0 / 3.0))
    rho_j = float(s["ETA_RHO"]) * float(s["RHO_AMB"])
    h_j = 1.0 + g / (g - 1.0) * float(s["P_EQ"]) / rho_j
    return float(s["GAMMA_JET"]) * math.sqrt(sigma * rho_j * h_j)


def apply_derived_ic_keys(s):
    """Settings-level consequences of the additive keys (in place)."""
    if s.get("NOZZLE_SIGMA") is not None:
        s["B0"] = nozzle_b0_from_sigma(s)
    return s


def init_block(s):
    """solvers/srhd3d_mpi_muscl.py:128-163 for one rank holding the whole domain."""
    rm = s["PHYSICS"] in ("rmhd", "grmhd")
    nv = 9 if rm else (15 if s.get("DISSIPATION_ENABLED", False) else 5)  # Pi, pi^ij, q^i at 5..14, zero
    nv += int(s.get("N_TRACERS", 0)) + int(s.get("N_THERMO", 0)) + int(s.get("N_CHEM", 0))
    ng = int(s["NG"])
    pr = np.zeros((nv, s["NX"] + 2 * ng, s["NY"] + 2 * ng, s["NZ"] + 2 * ng), dtype=np.float64)
    pr[0] = s["RHO_AMB"]
    pr[1] = s["VX_AMB"]
    pr[2] = s["VY_AMB"]
    pr[3] = s["VZ_AMB"]
    pr[4] = s["P_AMB"]
    if int(s.get("N_TRACERS", 0)) > 0:
        amb = s.get("TRACER_AMB_VALUES", [])
        for t in range(int(s["N_TRACERS"])):
            pr[int(s["TRACER_OFFSET"]) + t] = amb[t] if t < len(amb) else 0.0
    if int(s.get("N_THERMO", 0)) > 0:
        pr[int(s["THERMO_OFFSET"])] = float(s.get("TE_AMB", 0.0) or 0.0)
        pr[int(s["THERMO_OFFSET"]) + 1] = float(s.get("TI_AMB", 0.0) or 0.0)
    if int(s.get("N_CHEM", 0)) > 0:
        for m, key in enumerate(("CHEM_X_HII_AMB", "CHEM_X_HEII_AMB", "CHEM_X_HEIII_AMB")):
            pr[int(s["CHEM_OFFSET"]) + m] = float(s.get(key, 0.0))
    return pr


def add_density_noise(s, pr):
    a = float(s.get("IC_RHO_NOISE", 0.0) or 0.0)
    if a <= 0.0:
        return pr
    ng = int(s["NG"])
    noise = np.random.default_rng(int(s.get("IC_SEED", 12345))).uniform(-a, a, (s["NX"], s["NY"], s["NZ"]))
    pr[0, ng:-ng, ng:-ng, ng:-ng] = s["RHO_AMB"] * (1.0 + noise)
    return pr


def vector_potential_field(shape, kpeak, seed, amp):
    """Divergence-free periodic perturbation dB = curl A (2nd-order centred) with rms |dB| = amp."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    kz = np.fft.rfftfreq(nz) * nz
    K = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    w = np.exp(-((K - kpeak) ** 2) / 8.0)
    w[0, 0, 0] = 0.0
    A = []
    for _ in range(3):
        modes = (rng.normal(size=K.shape) + 1j * rng.normal(size=K.shape)) * w
        A.append(np.fft.irfftn(modes, s=shape, axes=(0, 1, 2)))
    Ax, Ay, Az = A

    def d(f, axis, h):
        return (np.roll(f, -1, axis) - np.roll(f, 1, axis)) / (2.0 * h)

    hx, hy, hz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    bx = d(Az, 1, hy) - d(Ay, 2, hz)
    by = d(Ax, 2, hz) - d(Az, 0, hx)
    bz = d(Ay, 0, hx) - d(Ax, 1, hy)
    rms = math.sqrt(float(np.mean(bx ** 2 + by ** 2 + bz ** 2)))
    f = amp / rms if rms > 0 else 0.0
    return bx * f, by * f, bz * f


def add_magnetic_field(s, pr):
    if s["PHYSICS"] not in ("rmhd", "grmhd"):
        return pr
    bamb = s.get("B_AMB")
    if bamb:
        for m in range(3):
            pr[5 + m] += float(bamb[m])
    amp = float(s.get("B_PERT_AMP", 0.0) or 0.0)
    if amp > 0.0:
        ng = int(s["NG"])
        bx, by, bz = vector_potential_field((s["NX"], s["NY"], s["NZ"]), float(s.get("B_PERT_KPEAK", 8.0)),
                                            int(s.get("B_PERT_SEED", 54321)), amp)
        sl = (slice(ng, -ng),) * 3
        pr[5][sl] += bx
        pr[6][sl] += by
        pr[7][sl] += bz
    return pr


def initial_state(s):
    """init_block + the additive IC keys, ghosts still at ambient (the first BC fill sets them)."""
    apply_derived_ic_keys(s)
    pr = init_block(s)
    add_density_noise(s, pr)
    add_magnetic_field(s, pr)
    return pr
