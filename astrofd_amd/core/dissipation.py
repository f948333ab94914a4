"""Israel-Stewart dissipation with the name of core/dissipation.py, on the MI355X library."""
from . import _ctx
from ..domain import diss_from_settings


def apply_causal_dissipation(pr, dt, dx, dy, dz, cfg):
    """core/dissipation.py:5-215: upwind advection (DISSIPATION_ADVECT) and IMEX relaxation of the
    bulk pressure, shear stress and heat flux held in fields 5..14, on interior cells; returns pr
    (updated in place), unchanged when the option is off, for RMHD or with fewer than 15 fields."""
    if not cfg.get("DISSIPATION_ENABLED", False) or cfg.get("PHYSICS") not in ("hydro", "grhd"):
        return pr
    if pr.shape[0] < 15:
        return pr
    d = _ctx.source_domain(pr.shape, int(cfg.get("NG", 2)), dxyz=(dx, dy, dz))
    d.set_prim(pr)
    d.causal_dissipation(dt, diss_from_settings(cfg))
    pr[...] = d.get_prim()
    return pr
