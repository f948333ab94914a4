"""Nozzle inflow with the names of core/nozzle.py, on the MI355X library."""
import ctypes
import math

from .. import _lib as L
from ..domain import nozzle_from_settings
from . import _ctx
from .boundary import _physics


def nozzle_weight(r, radius, shear_thick, profile):
    """core/nozzle.py:5-25: the profile weight s(r) in [0, 1]."""
    if radius <= 0.0:
        return 0.0
    if profile in ("taper", "smooth"):  # raised cosine
        return 0.0 if r >= radius else 0.5 * (1.0 + math.cos(math.pi * r / radius))
    if profile in ("parabolic", "parabola"):
        if r >= radius:
            return 0.0
        x = r / radius
        return max(0.0, 1.0 - x * x)
    if shear_thick > 0.0:  # top hat with a tanh shear layer
        return 0.5 * (1.0 - math.tanh((r - radius) / shear_thick))
    return 1.0 if r <= radius else 0.0


def apply_nozzle_left_x(pr, dx, dy, dz, ny_loc, nz_loc, y0, z0, rng, cfg):
    """core/nozzle.py:27-92, in place on pr.  `rng` (numpy Generator) supplies the perturbation
    draws in the reference's order when NOZZLE_PERTURB is not "none"."""
    phys = _physics(pr)
    d = _ctx.domain(phys, pr.shape, (dx, dy, dz), ng=int(cfg["NG"]))
    if ny_loc != pr.shape[2] - 2 * int(cfg["NG"]) or nz_loc != pr.shape[3] - 2 * int(cfg["NG"]):
        raise ValueError("ny_loc / nz_loc must be the interior y / z sizes of pr")
    n = nozzle_from_settings(dict(cfg, JET_CENTER=[0.0, float(y0), float(z0)]))
    L.check(d.lib.afd_set_nozzle(d.h, ctypes.byref(n)), "afd_set_nozzle")
    d.nozzle_set = True
    d.set_prim(pr)
    d.apply_nozzle_left_x(rng)
    pr[...] = d.get_prim()
