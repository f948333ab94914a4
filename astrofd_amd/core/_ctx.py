"""Device contexts behind the reference-named modules, cached per configuration and shape."""
import numpy as np

from .. import _lib as L
from ..domain import Domain, params_from_settings
from ..settings import finalize

_SETTINGS = {"hydro": None, "rmhd": None}
_CACHE = {}


def configure(physics, params):
    """Keep the settings dict for `physics` ("hydro" or "rmhd"), as the reference's configure()
    freezes its module globals; drops contexts made under an earlier configuration."""
    s = finalize(dict(params)) if "N_PASSIVE" not in params else dict(params)
    _SETTINGS[physics] = s
    for key in [k for k in _CACHE if k[0] == physics]:
        _CACHE.pop(key).close()


def settings(physics):
    s = _SETTINGS[physics]
    if s is None:
        raise RuntimeError(f"configure(params) must run before the {physics} core is used "
                           "(core/srhd_core.py:18, core/rmhd_core.py:21)")
    return s


def domain(physics, full_shape, dxyz=(1.0, 1.0, 1.0), rk_order=2, ng=None, gamma=None, gr=None):
    """A context for arrays of shape (nv, NX+2NG, NY+2NG, NZ+2NG).  `gr`: a dict of GR_METRIC,
    GR_MASS, ORTHONORMAL_FLUX and the global x offset "offs_x" for the GR wrappers."""
    s = settings(physics)
    ng = int(s["NG"] if ng is None else ng)
    nv, fx, fy, fz = (int(v) for v in full_shape)
    grkey = tuple(sorted(gr.items())) if gr else None
    key = (physics, nv, fx, fy, fz, ng, tuple(float(d) for d in dxyz), int(rk_order), gamma, grkey)
    d = _CACHE.get(key)
    if d is None:
        s2 = dict(s, PHYSICS=physics)
        over = dict(nx=fx - 2 * ng, ny=fy - 2 * ng, nz=fz - 2 * ng, ng=ng, dx=float(dxyz[0]), dy=float(dxyz[1]),
                    dz=float(dxyz[2]), rk_order=int(rk_order), nv=nv)
        nb = 9 if physics == "rmhd" else 5
        po = int(s2.get("PASSIVE_OFFSET", nb)) if physics == "hydro" else nb
        po = po if nv >= po else nb  # arrays without the dissipative block
        over["passive_offset"] = po
        over["n_passive"] = nv - po
        if gamma is not None:
            over["gamma"] = float(gamma)
        # NX..NZ and Lx..Lz only size the settings-derived dx (the reference's configure() reads
        # neither); the overrides set the grid
        s2.update(NX=max(fx - 2 * ng, 1), NY=max(fy - 2 * ng, 1), NZ=max(fz - 2 * ng, 1))
        for key in ("Lx", "Ly", "Lz"):
            s2.setdefault(key, 1.0)
        if gr:
            s2.update(PHYSICS="grmhd" if physics == "rmhd" else "grhd", GR_METRIC=gr["GR_METRIC"],
                      GR_MASS=gr["GR_MASS"], ORTHONORMAL_FLUX=gr["ORTHONORMAL_FLUX"])
            over["gr_x0"] = int(gr.get("offs_x", 0))
        d = Domain(params_from_settings(s2, **over))
        _CACHE[key] = d
    return d


def source_domain(full_shape, ng, two_temp=None, dxyz=(1.0, 1.0, 1.0)):
    """A context for the post-step sources on arrays of `full_shape`.  They read their constants
    from the cfg of each call (core/source_terms.py, core/chemistry.py), so no configure() is
    needed; `two_temp` = (THERMO_OFFSET, TEI_TAU) or None."""
    nv, fx, fy, fz = (int(v) for v in full_shape)
    key = ("sources", nv, fx, fy, fz, int(ng), two_temp, tuple(float(d) for d in dxyz))
    d = _CACHE.get(key)
    if d is None:
        nx, ny, nz = max(fx - 2 * ng, 1), max(fy - 2 * ng, 1), max(fz - 2 * ng, 1)
        s = dict(PHYSICS="hydro", NX=nx, NY=ny, NZ=nz, NG=int(ng), Lx=1.0, Ly=1.0, Lz=1.0)
        over = dict(nv=nv, n_passive=nv - 5, passive_offset=5, dx=float(dxyz[0]), dy=float(dxyz[1]),
                    dz=float(dxyz[2]))
        if two_temp is not None:
            over.update(two_temperature=1, thermo_offset=int(two_temp[0]), tei_tau=float(two_temp[1]))
        d = Domain(params_from_settings(s, **over))
        _CACHE[key] = d
    return d


def pointwise(physics, gamma=None):
    """A small context for the pointwise conversions (its grid is unused)."""
    return domain(physics, (9 if physics == "rmhd" else 5, 8, 8, 8), ng=2, gamma=gamma)


def as_batch(args, nb):
    """Scalars or arrays of one shape -> (n x nb float64 array, shape)."""
    arrs = np.broadcast_arrays(*[np.asarray(a, dtype=np.float64) for a in args])
    shape = arrs[0].shape
    return np.ascontiguousarray(np.stack([a.ravel() for a in arrs], axis=1)), shape


def unbatch(out, shape):
    if shape == ():
        return tuple(float(v) for v in out[0])
    return tuple(out[:, m].reshape(shape) for m in range(out.shape[1]))


def close_all():
    for d in _CACHE.values():
        d.close()
    _CACHE.clear()


__all__ = ["configure", "settings", "domain", "source_domain", "pointwise", "as_batch", "unbatch", "close_all", "L"]
