"""RMHD operators with the names and conventions of core/rmhd_core.py, on the MI355X library.

configure (:21), prim_to_cons_rmhd (:138), cons_to_prim_rmhd (:249), compute_rhs_rmhd
(MUSCL, :749), compute_rhs_ppm (:1280), compute_rhs_weno (:1440), step_ssprk2 (:1559),
step_ssprk3 (:1609).  Gamma is the configured GAMMA (the RMHD helpers pass it explicitly).
"""
from . import _ctx


def configure(params):
    """core/rmhd_core.py:21-50: freeze the settings used by the calls below."""
    _ctx.configure("rmhd", params)


def prim_to_cons_rmhd(rho, vx, vy, vz, p, Bx, By, Bz, psi, gamma=None):
    """core/rmhd_core.py:138-168 -> (D, Sx, Sy, Sz, tau, Bx, By, Bz, psi)."""
    q, shape = _ctx.as_batch((rho, vx, vy, vz, p, Bx, By, Bz, psi), 9)
    return _ctx.unbatch(_ctx.pointwise("rmhd", gamma).p2c(q), shape)


def cons_to_prim_rmhd(D, Sx, Sy, Sz, tau, Bx, By, Bz, psi):
    """core/rmhd_core.py:249-341 -> (rho, vx, vy, vz, p, Bx, By, Bz, psi)."""
    u, shape = _ctx.as_batch((D, Sx, Sy, Sz, tau, Bx, By, Bz, psi), 9)
    return _ctx.unbatch(_ctx.pointwise("rmhd").c2p(u), shape)


def _rhs(pr, recon, dx, dy, dz):
    s = _ctx.settings("rmhd")
    if str(s.get("RECON", "muscl")).lower() != recon:
        _ctx.configure("rmhd", dict(s, RECON=recon))
    d = _ctx.domain("rmhd", pr.shape, (dx, dy, dz))
    d.set_prim(pr)
    return d.compute_rhs()


def compute_rhs_rmhd(pr, nx, ny, nz, dx, dy, dz):
    """core/rmhd_core.py:749-1277 (MUSCL): rhs (9 + passives, nx, ny, nz)."""
    return _rhs(pr, "muscl", dx, dy, dz)


def compute_rhs_ppm(pr, nx, ny, nz, dx, dy, dz):
    """core/rmhd_core.py:1280-1437."""
    return _rhs(pr, "ppm", dx, dy, dz)


def compute_rhs_weno(pr, nx, ny, nz, dx, dy, dz):
    """core/rmhd_core.py:1440-1557."""
    return _rhs(pr, "weno", dx, dy, dz)


def _step(pr, dx, dy, dz, dt, order):
    d = _ctx.domain("rmhd", pr.shape, (dx, dy, dz), rk_order=order)
    d.set_prim(pr)
    d.step(dt)
    return d.get_prim()


def step_ssprk2(pr, dx, dy, dz, dt):
    """core/rmhd_core.py:1559-1607: returns the new primitive array (pr is not modified)."""
    return _step(pr, dx, dy, dz, dt, 2)


def step_ssprk3(pr, dx, dy, dz, dt):
    """core/rmhd_core.py:1609-1677."""
    return _step(pr, dx, dy, dz, dt, 3)
