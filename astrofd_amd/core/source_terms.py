"""Post-step sources with the names of core/source_terms.py, on the MI355X library.

Each call copies `pr` to a device context sized for it, runs the source kernel over the interior
and copies the result back into `pr`; the constants come from `cfg`, as in the reference."""
from . import _ctx


def apply_cooling_heating(pr, dt, cfg):
    """core/source_terms.py:5-28: p <- min(max(p + dt (HEATING_RATE - COOLING_LAMBDA rho p), 1e-12),
    P_MAX) on interior cells; returns pr (updated in place), unchanged when the option is off."""
    if not cfg.get("COOLING_ENABLED", False):
        return pr
    cool, heat = float(cfg.get("COOLING_LAMBDA", 0.0)), float(cfg.get("HEATING_RATE", 0.0))
    if cool == 0.0 and heat == 0.0:
        return pr
    d = _ctx.source_domain(pr.shape, int(cfg.get("NG", 2)))
    d.set_prim(pr)
    d.cooling_heating(dt, cool, heat, float(cfg.get("P_MAX", 1.0)))
    pr[...] = d.get_prim()
    return pr


def apply_two_temperature(pr, dt, cfg):
    """core/source_terms.py:30-55: Te, Ti relax towards T = p / max(rho, 1e-12) on interior
    cells; returns pr (updated in place), unchanged when the option is off."""
    if not cfg.get("TWO_TEMPERATURE", False):
        return pr
    off = int(cfg.get("THERMO_OFFSET", 0))
    tau = float(cfg.get("TEI_TAU", 0.5))
    if pr.shape[0] < off + 2 or tau <= 0.0:
        return pr
    d = _ctx.source_domain(pr.shape, int(cfg.get("NG", 2)), two_temp=(off, tau))
    d.set_prim(pr)
    d.two_temperature(dt)
    pr[...] = d.get_prim()
    return pr
