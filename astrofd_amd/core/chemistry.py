"""H/He ion chemistry with the name of core/chemistry.py, on the MI355X library."""
from . import _ctx
from ..domain import chem_from_settings


def apply_ion_chemistry(pr, dt, cfg):
    """core/chemistry.py:43-139: the IMEX update of xHII, xHeII, xHeIII at CHEM_OFFSET and its
    pressure coupling on interior cells; returns pr (updated in place), unchanged when the option
    is off, N_CHEM < 3 or the array has no room for the fractions."""
    if not cfg.get("CHEMISTRY_ENABLED", False) or int(cfg.get("N_CHEM", 0)) < 3:
        return pr
    if pr.shape[0] <= int(cfg.get("CHEM_OFFSET", 0)) + 2:
        return pr
    d = _ctx.source_domain(pr.shape, int(cfg.get("NG", 2)))
    d.set_prim(pr)
    d.ion_chemistry(dt, chem_from_settings(cfg))
    pr[...] = d.get_prim()
    return pr
