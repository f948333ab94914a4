"""GRHD on a fixed metric with the names of core/grhd_core.py, on the MI355X library.

configure (:13-18), compute_rhs_grhd(pr, dx, dy, dz, offs_x, ng) (:20-75), step_ssprk2 and
step_ssprk3 (pr, dx, dy, dz, dt, offs_x, ng).  The flat settings come from srhd_core.configure(), as in
the reference; the GR keys from this module's configure().  offs_x is the global x index of the
array's first interior cell (0 when one array holds the whole domain).
"""
from . import _ctx

GR_METRIC = "minkowski"
GR_MASS = 1.0
ORTHONORMAL_FLUX = True
GR_SPIN = 0.0


def configure(params):
    """core/grhd_core.py:13-18."""
    global GR_METRIC, GR_MASS, ORTHONORMAL_FLUX, GR_SPIN
    GR_METRIC = str(params.get("GR_METRIC", "minkowski")).lower()
    GR_MASS = float(params.get("GR_MASS", GR_MASS))
    GR_SPIN = float(params.get("GR_SPIN", GR_SPIN))
    ORTHONORMAL_FLUX = bool(params.get("ORTHONORMAL_FLUX", ORTHONORMAL_FLUX))


def _gr(offs_x):
    return dict(GR_METRIC=GR_METRIC, GR_MASS=GR_MASS, ORTHONORMAL_FLUX=ORTHONORMAL_FLUX, offs_x=int(offs_x))


def compute_rhs_grhd(pr, dx, dy, dz, offs_x, ng):
    """core/grhd_core.py:20-75: the flat RHS of the configured reconstruction with the flux velocities
    divided by the Kerr-Schild lapse (ORTHONORMAL_FLUX, Schwarzschild or Kerr-Schild), plus the
    lapse-gradient sources for every metric but Minkowski."""
    d = _ctx.domain("hydro", pr.shape, (dx, dy, dz), ng=ng, gr=_gr(offs_x))
    d.set_prim(pr)
    return d.compute_rhs()


def _step(pr, dx, dy, dz, dt, offs_x, ng, order):
    d = _ctx.domain("hydro", pr.shape, (dx, dy, dz), rk_order=order, ng=ng, gr=_gr(offs_x))
    d.set_prim(pr)
    d.step(dt)
    return d.get_prim()


def step_ssprk2(pr, dx, dy, dz, dt, offs_x, ng):
    """core/grhd_core.py: returns the new primitive array (pr is not modified)."""
    return _step(pr, dx, dy, dz, dt, offs_x, ng, 2)


def step_ssprk3(pr, dx, dy, dz, dt, offs_x, ng):
    """core/grhd_core.py."""
    return _step(pr, dx, dy, dz, dt, offs_x, ng, 3)
