"""SRHD operators with the names and conventions of core/srhd_core.py, on the MI355X library.

configure (:18), prim_to_cons (:47), cons_to_prim (:62), compute_rhs_muscl (:469),
compute_rhs_ppm (:927), compute_rhs_weno (:1167), max_char_speed (:1268),
step_ssprk2 (:1286), step_ssprk3 (:1334).  The pointwise conversions take scalars (or
numpy arrays of one shape, evaluated element by element) and return a tuple.
"""
import numpy as np

from . import _ctx

GAMMA = 5.0 / 3.0


def configure(params):
    """core/srhd_core.py:18-41: freeze the settings used by the calls below."""
    _ctx.configure("hydro", params)


def prim_to_cons(rho, vx, vy, vz, p, gamma=GAMMA):
    """core/srhd_core.py:47-59 -> (D, Sx, Sy, Sz, tau)."""
    q, shape = _ctx.as_batch((rho, vx, vy, vz, p), 5)
    return _ctx.unbatch(_ctx.pointwise("hydro", gamma).p2c(q), shape)


def cons_to_prim(D, Sx, Sy, Sz, tau, gamma=GAMMA):
    """core/srhd_core.py:62-131 -> (rho, vx, vy, vz, p)."""
    u, shape = _ctx.as_batch((D, Sx, Sy, Sz, tau), 5)
    return _ctx.unbatch(_ctx.pointwise("hydro", gamma).c2p(u), shape)


def _rhs(pr, recon, dx, dy, dz):
    s = _ctx.settings("hydro")
    if str(s.get("RECON", "muscl")).lower() != recon:
        _ctx.configure("hydro", dict(s, RECON=recon))
    d = _ctx.domain("hydro", pr.shape, (dx, dy, dz))
    d.set_prim(pr)
    return d.compute_rhs()


def compute_rhs_muscl(pr, nx, ny, nz, dx, dy, dz):
    """core/srhd_core.py:469-924: rhs (5 + passives, nx, ny, nz); nx..nz are the full dims of pr."""
    return _rhs(pr, "muscl", dx, dy, dz)


def compute_rhs_ppm(pr, nx, ny, nz, dx, dy, dz):
    """core/srhd_core.py:927-1162."""
    return _rhs(pr, "ppm", dx, dy, dz)


def compute_rhs_weno(pr, nx, ny, nz, dx, dy, dz):
    """core/srhd_core.py:1167-1262."""
    return _rhs(pr, "weno", dx, dy, dz)


def max_char_speed(pr, nx, ny, nz):
    """core/srhd_core.py:1268-1281: max |lambda| over [1, n-1)^3."""
    d = _ctx.domain("hydro", pr.shape)
    d.set_prim(pr)
    return d.max_char_speed()


def _step(pr, dx, dy, dz, dt, order):
    d = _ctx.domain("hydro", pr.shape, (dx, dy, dz), rk_order=order)
    d.set_prim(pr)
    d.step(dt)
    return d.get_prim()


def step_ssprk2(pr, dx, dy, dz, dt):
    """core/srhd_core.py:1286-1332: returns the new primitive array (pr is not modified)."""
    return _step(pr, dx, dy, dz, dt, 2)


def step_ssprk3(pr, dx, dy, dz, dt):
    """core/srhd_core.py:1334-1400."""
    return _step(pr, dx, dy, dz, dt, 3)
