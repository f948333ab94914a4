"""Seeded primitive states used by the golden-vector script and the larger GPU tests."""
import numpy as np


def srhd_state(shape, seed, npass=0, jet=True):
    """Ambient rho ~ 1 +- 30 %, |v| <= 0.05, p ~ 0.01, with a W ~ 7 jet core and a quiescent row."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    pr = np.zeros((5 + npass, nx, ny, nz))
    pr[0] = 1.0 + 0.3 * rng.uniform(-1, 1, shape)
    pr[1] = 0.05 * rng.uniform(-1, 1, shape)
    pr[2] = 0.05 * rng.uniform(-1, 1, shape)
    pr[3] = 0.05 * rng.uniform(-1, 1, shape)
    pr[4] = 0.01 * (1.0 + 0.5 * rng.uniform(-1, 1, shape))
    if jet:
        sl = (slice(0, nx // 2 + 1), slice(ny // 3, 2 * ny // 3 + 1), slice(nz // 3, 2 * nz // 3 + 1))
        pr[0][sl] = 0.01 * (1.0 + 0.2 * rng.uniform(-1, 1, pr[0][sl].shape))
        pr[1][sl] = 0.99 - 0.01 * rng.uniform(0, 1, pr[1][sl].shape)
        pr[4][sl] = 0.01 * (1.0 + 0.3 * rng.uniform(-1, 1, pr[4][sl].shape))
    pr[1:4, 0, 0, :] = 0.0
    for t in range(npass):
        pr[5 + t] = rng.uniform(0.0, 1.0, shape)
    return pr


def rmhd_state(shape, seed, npass=0, jet=True):
    """srhd_state plus Bx ~ 0.1, By, Bz ~ +-0.05, psi ~ 1e-3 and one z-plane with Bx = 0."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    pr = np.zeros((9 + npass, nx, ny, nz))
    pr[0:5] = srhd_state(shape, seed + 1000, 0, jet)
    pr[5] = 0.1 + 0.02 * rng.uniform(-1, 1, shape)
    pr[6] = 0.05 * rng.uniform(-1, 1, shape)
    pr[7] = 0.05 * rng.uniform(-1, 1, shape)
    pr[8] = 1e-3 * rng.uniform(-1, 1, shape)
    pr[5, :, :, 1] = 0.0
    for t in range(npass):
        pr[9 + t] = rng.uniform(0.0, 1.0, shape)
    return pr
