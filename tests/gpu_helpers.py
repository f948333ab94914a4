"""Helpers shared by the GPU tests: build a Domain for a reference-layout array."""
import numpy as np

from astrofd_amd import _lib as L
from astrofd_amd.domain import Domain, limiter_id, recon_id, riemann_id


def make_domain(physics, pr_shape, ng, dxyz=(0.1, 0.1, 0.1), gamma=None, p_max=2.0, v_max=0.9999, limiter="mc",
                recon="muscl", riemann="hlle", rk_order=2, n_passive=0, n_resist=0, resistivity=0.0,
                precision="fp64", glm_ch=1.0, glm_cp=0.1, bc_y="periodic", bc_z="periodic",
                two_temperature=False, thermo_offset=None, tei_tau=0.1, c2p_guess="plain", passive_offset=None,
                gr_metric="minkowski", gr_mass=1.0, orthonormal_flux=True, gr_gamma=None, face_path=0):
    rm = physics == "rmhd"
    nb = 9 if rm else 5
    p = L.Params()
    p.physics = L.AFD_PHYS_RMHD if rm else L.AFD_PHYS_HYDRO
    p.nv = pr_shape[0]
    p.n_passive = n_passive
    p.passive_offset = nb if passive_offset is None else passive_offset
    p.nx, p.ny, p.nz = (pr_shape[1] - 2 * ng, pr_shape[2] - 2 * ng, pr_shape[3] - 2 * ng)
    p.ng = ng
    p.dx, p.dy, p.dz = dxyz
    p.gamma = (gamma if gamma is not None else 5.0 / 3.0) if rm else 5.0 / 3.0
    p.p_max, p.v_max, p.glm_ch, p.glm_cp = p_max, v_max, glm_ch, glm_cp
    p.limiter_id = limiter_id(limiter)
    p.recon_id = recon_id(recon)
    p.riemann_id = riemann_id(physics, riemann)
    p.rk_order = rk_order
    p.n_resist = n_resist
    p.resistivity = resistivity
    p.precision = L.AFD_PREC_MIXED if precision == "mixed" else L.AFD_PREC_FP64
    p.c2p_guess = {"learned": L.AFD_C2P_LEARNED, "previous": L.AFD_C2P_PREVIOUS}.get(c2p_guess, L.AFD_C2P_PLAIN)
    p.bc_y = L.AFD_BC_OUTFLOW if bc_y == "outflow" else L.AFD_BC_PERIODIC
    p.bc_z = L.AFD_BC_OUTFLOW if bc_z == "outflow" else L.AFD_BC_PERIODIC
    p.two_temperature = int(two_temperature)
    p.thermo_offset = thermo_offset if thermo_offset is not None else nb
    p.tei_tau = tei_tau
    p.device = 0
    p.face_path = face_path
    from astrofd_amd.domain import gr_metric_id
    p.gr_metric = gr_metric_id(gr_metric)
    p.gr_orthonormal = int(bool(orthonormal_flux))
    p.gr_mass = gr_mass
    p.gr_gamma = gr_gamma if gr_gamma is not None else p.gamma
    return Domain(p)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b).max(initial=0.0)
    s = np.abs(b).max(initial=0.0)
    return d / s if s > 0 else d


def log(*a):
    print(*a, flush=True)
