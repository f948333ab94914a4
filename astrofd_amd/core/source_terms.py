"""Post-step sources with the names of core/source_terms.py, on the MI355X library."""
from . import _ctx
from .boundary import _physics


def apply_two_temperature(pr, dt, cfg):
    """core/source_terms.py:30-55: Te, Ti relax towards T = p / max(rho, 1e-12) on interior
    cells; returns pr (updated in place), unchanged when the option is off."""
    if not cfg.get("TWO_TEMPERATURE", False):
        return pr
    off = int(cfg.get("THERMO_OFFSET", 0))
    if pr.shape[0] < off + 2 or float(cfg.get("TEI_TAU", 0.5)) <= 0.0:
        return pr
    d = _ctx.domain(_physics(pr), pr.shape, ng=int(cfg.get("NG", 2)))
    if d.params.thermo_offset != off or d.params.tei_tau != float(cfg.get("TEI_TAU", 0.5)) \
            or not d.params.two_temperature:
        raise ValueError("TWO_TEMPERATURE / THERMO_OFFSET / TEI_TAU differ from the configured settings")
    d.set_prim(pr)
    d.two_temperature(dt)
    pr[...] = d.get_prim()
    return pr


def apply_cooling_heating(pr, dt, cfg):
    """core/source_terms.py:5-28.  Not on this path: a no-op when disabled, an error otherwise."""
    if not cfg.get("COOLING_ENABLED", False):
        return pr
    if float(cfg.get("COOLING_LAMBDA", 0.0)) == 0.0 and float(cfg.get("HEATING_RATE", 0.0)) == 0.0:
        return pr
    raise NotImplementedError("cooling/heating is not on the MI355X path")
