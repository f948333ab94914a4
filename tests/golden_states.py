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


def smooth_state(shape, seed, physics="hydro", vamp=0.2, npass=0):
    """Well-conditioned state: rho, p, v (and B, psi) as a few random Fourier modes around
    rho = 1, p = 0.05, |v| <= vamp, B ~ 0.1 (RMHD).  Used where a step is compared across
    implementations: rough states make the reference scheme diverge chaotically within one step."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.arange(nz)[None, None, :] / nz

    def field(amp):
        f = np.zeros(shape)
        for _ in range(4):
            k = rng.integers(1, 3, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            f += np.sin(2 * np.pi * k[0] * x + ph[0]) * np.sin(2 * np.pi * k[1] * y + ph[1]) * \
                np.sin(2 * np.pi * k[2] * z + ph[2])
        return amp * f / 4.0

    nb = 9 if physics == "rmhd" else 5
    pr = np.zeros((nb + npass,) + tuple(shape))
    pr[0] = 1.0 + field(0.3)
    pr[1] = field(vamp)
    pr[2] = field(vamp)
    pr[3] = field(vamp)
    pr[4] = 0.05 * (1.0 + field(0.5))
    if physics == "rmhd":
        pr[5] = 0.1 + field(0.05)
        pr[6] = field(0.05)
        pr[7] = field(0.05)
        pr[8] = field(1e-3)
    for t in range(npass):
        pr[nb + t] = 0.5 + field(0.4)
    return pr


def source_state(shape, seed, nv, chem_off=None, bad=True):
    """rho and p log-uniform over wide ranges (T = p / rho * CHEM_TUNIT_K from ~10 K to ~1e12 K),
    velocities and fields small, ion fractions in [-0.2, 1.2] so the clamps act, and (bad=True)
    a few cells with rho <= 0 or p <= 0, which chemistry skips.  The other variables are zero
    (the sources do not read them), which keeps the fixture small."""
    rng = np.random.default_rng(seed)
    pr = np.zeros((nv,) + shape)
    pr[0] = 10.0 ** rng.uniform(-3, 1, shape)
    pr[4] = 10.0 ** rng.uniform(-4, 3, shape)
    if chem_off is not None:
        pr[chem_off:chem_off + 3] = rng.uniform(-0.2, 1.2, (3,) + shape)
    if bad:
        flat = pr.reshape(nv, -1)
        n = flat.shape[1]
        cells = rng.choice(n, 12, replace=False)
        flat[0, cells[:4]] = [0.0, -1.0, 0.0, -1e-3]
        flat[4, cells[4:8]] = [0.0, -2.0, 0.0, -1e-6]
        flat[0, cells[8:10]] = 0.0
        flat[4, cells[10:12]] = 0.0
    return pr
