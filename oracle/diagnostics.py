"""CPU oracle of the run diagnostics (utils/diagnostics.py), numpy.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the GPU reductions.

  diagnostics(pr, ng, dx, dy, dz, rmhd, p_max, v_max):
      max Lorentz factor over the interior (utils/diagnostics.py:15-27), inlet flux
      sum of Sx dy dz on the plane i = NG (:31-43; the reference calls prim_to_cons*
      without gamma there, so their definition-time default 5/3 applies), and the
      centred div B max / rms / mean (:128-183)
  centerline(pr, ng, v_max): Gamma, rho, p, vx along x at the mid (j, k) (:59-126)

The p2c used for the inlet flux is the C oracle's (bitwise with the reference).
"""
import numpy as np

import oracle as O


def diagnostics(pr, ng, dx, dy, dz, rmhd, p_max=1.0, v_max=0.999):
    core = (slice(ng, -ng),) * 3
    v2 = pr[1][core] ** 2 + pr[2][core] ** 2 + pr[3][core] ** 2
    v2 = np.where(v2 >= 1.0, 1.0 - 1e-14, v2)
    max_gamma = float(np.max(1.0 / np.sqrt(1.0 - v2)))
    plane = pr[:, ng, ng:-ng, ng:-ng]
    nb = 9 if rmhd else 5
    states = plane[:nb].reshape(nb, -1).T
    P = O.make_params("rmhd" if rmhd else "hydro", gamma=5.0 / 3.0, p_max=p_max, v_max=v_max)
    U = O.rmhd_p2c(P, states) if rmhd else O.srhd_p2c(P, states)
    inlet = float(np.sum(U[:, 1] * dy * dz))
    out = dict(max_gamma=max_gamma, inlet_flux=inlet, divb_max=0.0, divb_rms=0.0, divb_mean=0.0)
    if rmhd:
        s = slice(ng, -ng)
        dBx = (pr[5, ng + 1:pr.shape[1] - ng + 1, s, s] - pr[5, ng - 1:pr.shape[1] - ng - 1, s, s]) / (2.0 * dx)
        dBy = (pr[6, s, ng + 1:pr.shape[2] - ng + 1, s] - pr[6, s, ng - 1:pr.shape[2] - ng - 1, s]) / (2.0 * dy)
        dBz = (pr[7, s, s, ng + 1:pr.shape[3] - ng + 1] - pr[7, s, s, ng - 1:pr.shape[3] - ng - 1]) / (2.0 * dz)
        d = dBx + dBy + dBz
        out.update(divb_max=float(np.max(np.abs(d))), divb_rms=float(np.sqrt(np.mean(d * d))),
                   divb_mean=float(np.mean(d)))
    return out


def centerline(pr, ng, v_max):
    j = ng + (pr.shape[2] - 2 * ng) // 2
    k = ng + (pr.shape[3] - 2 * ng) // 2
    sl = slice(ng, pr.shape[1] - ng)
    rho, vx, vy, vz, p = (pr[m, sl, j, k].copy() for m in range(5))
    v2 = vx * vx + vy * vy + vz * vz
    mask = v2 >= v_max * v_max
    if np.any(mask):
        fac = v_max / np.sqrt(v2[mask] + 1e-32)
        vx[mask] *= fac
        v2[mask] = v_max * v_max - 1e-14
    gamma = 1.0 / np.sqrt(1.0 - np.clip(v2, 0, 1 - 1e-14))
    return np.stack([gamma, rho, p, vx])
