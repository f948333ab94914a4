"""Device-resident domain: one libafd context holding the primitive state on the GPU.

`params_from_settings` restates what the reference's configure() functions
derive from the settings dict (core/srhd_core.py:18-41, core/rmhd_core.py:21-50)
and what the driver passes per call (grid spacing, RK order, NG).
"""
import ctypes

import numpy as np

from . import _lib as L

LIMITERS = {"minmod": 1, "vanleer": 2}  # anything else -> mc (0)


def limiter_id(name):
    return LIMITERS.get(str(name).lower(), 0)


def recon_id(name):
    name = str(name).lower()
    return 1 if name == "ppm" else (2 if name == "weno" else 0)


def riemann_id(physics, name):
    name = str(name).lower()
    if physics in ("rmhd", "grmhd"):  # core/rmhd_core.py:42-46: "hllc" falls through to HLLE
        return 2 if name == "hlld_full" else (1 if name == "hlld" else 0)
    return 1 if name == "hllc" else 0  # core/srhd_core.py:38-39


def n_resist_for(s, recon):
    """Applications of the resistive Laplacian per RHS.  Reference: once in MUSCL
    (core/rmhd_core.py:1265-1275), twice in PPM (:1414-1435), never in WENO (:1555)."""
    if not (s.get("RESISTIVE_ENABLED", False) and float(s.get("RESISTIVITY", 0.0)) > 0.0):
        return 0
    if str(s.get("RESISTIVE_MODE", "reference")).lower() == "once":
        return 1
    return {0: 1, 1: 2, 2: 0}[recon]


def params_from_settings(s, device=0, **override):
    phys = str(s.get("PHYSICS", "hydro")).lower()
    if phys not in ("hydro", "rmhd", "grhd", "grmhd"):
        raise NotImplementedError(f"PHYSICS={phys!r}: the path runs hydro, rmhd, grhd and grmhd")
    rm = phys in ("rmhd", "grmhd")
    nb = 9 if rm else 5
    rc = recon_id(s.get("RECON", "muscl"))
    p = L.Params()
    p.physics = L.AFD_PHYS_RMHD if rm else L.AFD_PHYS_HYDRO
    p.n_passive = int(s.get("N_PASSIVE", 0))
    p.passive_offset = int(s.get("PASSIVE_OFFSET", nb))  # 15 for SRHD with dissipation (fields 5..14)
    p.nv = p.passive_offset + p.n_passive
    p.nx, p.ny, p.nz, p.ng = int(s["NX"]), int(s["NY"]), int(s["NZ"]), int(s["NG"])
    p.dx = s["Lx"] / s["NX"]
    p.dy = s["Ly"] / s["NY"]
    p.dz = s["Lz"] / s["NZ"]
    # The SRHD helpers bind gamma=GAMMA at definition time (core/srhd_core.py:47, 62, 134), so the
    # reference SRHD step always computes with 5/3 whatever the config says; RMHD passes GAMMA.
    p.gamma = float(s.get("GAMMA", 5.0 / 3.0)) if rm else 5.0 / 3.0
    p.p_max = float(s.get("P_MAX", 1.0))
    p.v_max = float(s.get("V_MAX", 0.999))
    p.glm_ch = float(s.get("GLM_CH", 1.0))
    p.glm_cp = float(s.get("GLM_CP", 0.1))
    p.limiter_id = limiter_id(s.get("LIMITER", "mc"))
    p.recon_id = rc
    p.riemann_id = riemann_id(phys, s.get("RIEMANN", "hlle"))
    p.rk_order = 3 if int(s.get("RK_ORDER", 2)) == 3 else 2
    p.n_resist = n_resist_for(s, rc) if rm else 0
    p.resistivity = float(s.get("RESISTIVITY", 0.0))
    p.precision = L.AFD_PREC_MIXED if str(s.get("PRECISION", "fp64")).lower() == "mixed" else L.AFD_PREC_FP64
    p.c2p_guess = {"learned": L.AFD_C2P_LEARNED, "previous": L.AFD_C2P_PREVIOUS}.get(
        str(s.get("C2P_GUESS", "plain")).lower(), L.AFD_C2P_PLAIN)
    p.bc_y = L.AFD_BC_OUTFLOW if str(s.get("BC_Y", "periodic")).lower() == "outflow" else L.AFD_BC_PERIODIC
    p.bc_z = L.AFD_BC_OUTFLOW if str(s.get("BC_Z", "periodic")).lower() == "outflow" else L.AFD_BC_PERIODIC
    p.two_temperature = int(bool(s.get("TWO_TEMPERATURE", False)))
    p.thermo_offset = int(s.get("THERMO_OFFSET", nb))
    p.tei_tau = float(s.get("TEI_TAU", 0.5))
    p.device = device
    if phys in ("grhd", "grmhd"):  # core/grhd_core.py:13-18, core/grmhd_core.py:13-18
        p.gr_metric = gr_metric_id(s.get("GR_METRIC", "minkowski"))
        p.gr_orthonormal = int(bool(s.get("ORTHONORMAL_FLUX", True)))
        p.gr_mass = float(s.get("GR_MASS", 1.0))
        # the sources' enthalpy: srhd_core.GAMMA (the configured one) / rmhd_core.GAMMA
        p.gr_gamma = float(s.get("GAMMA", 5.0 / 3.0)) if not rm else p.gamma
    for k, v in override.items():
        setattr(p, k, v)
    return p


GR_METRICS = {"minkowski": 0, "schwarzschild": 1, "kerr-schild": 2}


def gr_metric_id(name):
    """core/grhd_core.py:23, 47, 56: Minkowski skips the wrapper; Schwarzschild and Kerr-Schild
    rescale the flux velocities; any other name takes the Kerr-Schild sources only (3)."""
    return GR_METRICS.get(str(name).lower(), 3)


PROFILES = {"top_hat": 0, "taper": 1, "smooth": 1, "parabolic": 2, "parabola": 2}
B_INITS = {"poloidal": 1, "toroidal": 2}


def nozzle_from_settings(s):
    """The nozzle parameters core/nozzle.py:27-92 reads from cfg (plus the additive B_AMB)."""
    n = L.Nozzle()
    jc = s["JET_CENTER"]
    n.y0, n.z0 = float(jc[1]), float(jc[2])
    n.jet_radius = float(s["JET_RADIUS"])
    n.shear_thick = float(s["SHEAR_THICK"])
    n.profile = PROFILES.get(str(s.get("NOZZLE_PROFILE", "top_hat")).lower(), 0)
    n.eta_rho, n.rho_amb, n.p_eq = float(s["ETA_RHO"]), float(s["RHO_AMB"]), float(s["P_EQ"])
    n.gamma_jet = float(s["GAMMA_JET"])
    n.vx_amb, n.vy_amb, n.vz_amb = float(s["VX_AMB"]), float(s["VY_AMB"]), float(s["VZ_AMB"])
    n.perturb = 0 if str(s.get("NOZZLE_PERTURB", "none")).lower() == "none" else 1
    n.turb_vamp, n.turb_pamp = float(s.get("TURB_VAMP", 0.0)), float(s.get("TURB_PAMP", 0.0))
    ntr = int(s.get("N_TRACERS", 0))
    if ntr > 8:
        raise ValueError("at most 8 tracers")
    n.n_tracers, n.tracer_offset = ntr, int(s.get("TRACER_OFFSET", 5))
    tn = list(s.get("TRACER_NOZZLE_VALUES") or [])
    ta = list(s.get("TRACER_AMB_VALUES") or [])
    for t in range(ntr):  # core/nozzle.py:78-79
        n.tracer_nozzle[t] = tn[t] if t < len(tn) else (1.0 if t == 0 else 0.0)
        n.tracer_amb[t] = ta[t] if t < len(ta) else 0.0
    n.two_temperature = int(bool(s.get("TWO_TEMPERATURE", False)))
    n.thermo_offset = int(s.get("THERMO_OFFSET", 0))
    te = s.get("TE_NOZZLE", s.get("TE_AMB", 0.0))
    ti = s.get("TI_NOZZLE", s.get("TI_AMB", 0.0))
    n.te_nozzle = float(te if te is not None else 0.0)
    n.ti_nozzle = float(ti if ti is not None else 0.0)
    n.b_init = B_INITS.get(str(s.get("B_INIT", "none")).lower(), 0)
    n.b0 = float(s.get("B0", 0.0))
    bamb = s.get("B_AMB") or [0.0, 0.0, 0.0]
    for m in range(3):
        n.b_amb[m] = float(bamb[m])
    if s.get("CHEMISTRY_ENABLED", False):  # core/nozzle.py:77-82
        n.chemistry, n.chem_offset = 1, int(s.get("CHEM_OFFSET", 0))
        for m, key in enumerate(("CHEM_X_HII_NOZZLE", "CHEM_X_HEII_NOZZLE", "CHEM_X_HEIII_NOZZLE")):
            v = s.get(key, 0.0)
            n.chem_nozzle[m] = float(v if v is not None else 0.0)
    return n


def chem_from_settings(s):
    """The constants core/chemistry.py:59-68 reads from the settings."""
    return L.Chem(offset=int(s.get("CHEM_OFFSET", 0)), X=float(s.get("CHEM_X", 0.76)), Y=float(s.get("CHEM_Y", 0.24)),
                  tunit_k=float(s.get("CHEM_TUNIT_K", 1.0e4)), rate_scale=float(s.get("CHEM_RATE_SCALE", 1.0)),
                  energy_scale=float(s.get("CHEM_ENERGY_SCALE", 1.0)), tmin_k=float(s.get("CHEM_TMIN_K", 1.0)),
                  gamma=float(s.get("GAMMA", 5.0 / 3.0)))


def diss_from_settings(s):
    """The constants core/dissipation.py:16-45 reads (a missing, None or zero RELAX_TAU_* falls
    back to RELAX_TAU, :19-21)."""
    base = s.get("RELAX_TAU", 0.1)

    def tau(key):
        return float(s.get(key, base) or base)
    return L.Diss(zeta=float(s.get("BULK_ZETA", 0.0)), eta=float(s.get("SHEAR_ETA", 0.0)),
                  kappa=float(s.get("HEAT_KAPPA", 0.0)), tau_bulk=tau("RELAX_TAU_BULK"),
                  tau_shear=tau("RELAX_TAU_SHEAR"), tau_heat=tau("RELAX_TAU_HEAT"),
                  cap_frac=float(s.get("DISSIPATION_CAP_FRAC", 0.5)), p_max=float(s.get("P_MAX", 1.0)),
                  advect=int(bool(s.get("DISSIPATION_ADVECT", True))))


def sources_from_settings(s):
    """The post-step sources the driver applies (solvers/srhd3d_mpi_muscl.py:324-326) and their
    settings gates: core/dissipation.py:10-13, core/source_terms.py:6-11, core/chemistry.py:48-53."""
    src = L.Sources()
    src.dissipation = int(bool(s.get("DISSIPATION_ENABLED", False)) and
                          str(s.get("PHYSICS", "hydro")).lower() in ("hydro", "grhd"))
    src.diss = diss_from_settings(s)
    src.cooling = int(bool(s.get("COOLING_ENABLED", False)))
    src.cool = float(s.get("COOLING_LAMBDA", 0.0))
    src.heat = float(s.get("HEATING_RATE", 0.0))
    src.p_cap = float(s.get("P_MAX", 1.0))
    src.chemistry = int(bool(s.get("CHEMISTRY_ENABLED", False)) and int(s.get("N_CHEM", 0)) >= 3)
    src.chem = chem_from_settings(s)
    return src


class Domain:
    """One device context: the primitive array `pr` of the reference, resident in HBM."""

    def __init__(self, params):
        self.lib = L.load()
        self.params = params
        self.nb = 9 if params.physics == L.AFD_PHYS_RMHD else 5
        self.shape = (params.nv, params.nx + 2 * params.ng, params.ny + 2 * params.ng, params.nz + 2 * params.ng)
        self.size = int(np.prod(self.shape))
        h = ctypes.c_void_p()
        L.check(self.lib.afd_create(ctypes.byref(params), ctypes.byref(h)), "afd_create")
        self.h = h
        self.nozzle_set = False
        self.sources = L.Sources()  # post-step sources beside two-temperature: none unless from settings

    @classmethod
    def from_settings(cls, s, device=0, **override):
        d = cls(params_from_settings(s, device, **override))
        d.sources = sources_from_settings(s)
        return d

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.afd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state
    def set_prim(self, pr):
        pr = L.f64c(pr)
        if pr.shape != self.shape:
            raise ValueError(f"pr shape {pr.shape} != {self.shape}")
        L.check(self.lib.afd_set_prim(self.h, L.dptr(pr), pr.size), "afd_set_prim")

    def get_prim(self):
        out = np.empty(self.shape)
        L.check(self.lib.afd_get_prim(self.h, L.dptr(out), out.size), "afd_get_prim")
        return out

    # ---- boundaries
    def set_nozzle(self, s):
        n = nozzle_from_settings(s)
        L.check(self.lib.afd_set_nozzle(self.h, ctypes.byref(n)), "afd_set_nozzle")
        self.nozzle_set = True

    def nozzle_draw_count(self):
        n = ctypes.c_int64()
        L.check(self.lib.afd_nozzle_draw_count(self.h, ctypes.byref(n)), "afd_nozzle_draw_count")
        return n.value

    def _draws(self, rng):
        n = self.nozzle_draw_count() if self.nozzle_set else 0
        if n == 0:
            return None, 0
        if rng is None:
            raise ValueError("a perturbed nozzle needs the numpy Generator of the run")
        return L.f64c(rng.random(n)), n

    def apply_bcs(self, rng=None):
        d, n = self._draws(rng)
        L.check(self.lib.afd_apply_bcs(self.h, L.dptr(d) if d is not None else None, n), "afd_apply_bcs")

    def apply_periodic_yz(self):
        L.check(self.lib.afd_apply_periodic_yz(self.h), "afd_apply_periodic_yz")

    def apply_outflow_right_x(self):
        L.check(self.lib.afd_apply_outflow_right_x(self.h), "afd_apply_outflow_right_x")

    def apply_nozzle_left_x(self, rng=None):
        d, n = self._draws(rng)
        L.check(self.lib.afd_apply_nozzle_left_x(self.h, L.dptr(d) if d is not None else None, n),
                "afd_apply_nozzle_left_x")

    # ---- operators
    def max_char_speed(self):
        a = ctypes.c_double()
        L.check(self.lib.afd_max_char_speed(self.h, ctypes.byref(a)), "afd_max_char_speed")
        return a.value

    def step(self, dt):
        L.check(self.lib.afd_step(self.h, float(dt)), "afd_step")

    def compute_rhs(self):
        out = np.empty((self.nb + self.params.n_passive,) + self.shape[1:])
        L.check(self.lib.afd_compute_rhs(self.h, L.dptr(out), out.size), "afd_compute_rhs")
        return out

    def two_temperature(self, dt):
        L.check(self.lib.afd_two_temperature(self.h, float(dt)), "afd_two_temperature")

    def cooling_heating(self, dt, cool, heat, p_cap=1.0):
        L.check(self.lib.afd_cooling_heating(self.h, float(dt), float(cool), float(heat), float(p_cap)),
                "afd_cooling_heating")

    def ion_chemistry(self, dt, chem):
        """`chem`: an L.Chem, or a settings dict (chem_from_settings)."""
        if isinstance(chem, dict):
            chem = chem_from_settings(chem)
        L.check(self.lib.afd_ion_chemistry(self.h, float(dt), ctypes.byref(chem)), "afd_ion_chemistry")

    def causal_dissipation(self, dt, diss):
        """`diss`: an L.Diss, or a settings dict (diss_from_settings)."""
        if isinstance(diss, dict):
            diss = diss_from_settings(diss)
        L.check(self.lib.afd_causal_dissipation(self.h, float(dt), ctypes.byref(diss)), "afd_causal_dissipation")

    def post_step_sources(self, dt, sources=None):
        """Dissipation, then cooling/heating, two-temperature and chemistry in one pass (the
        driver's order)."""
        src = self.sources if sources is None else sources
        L.check(self.lib.afd_post_step_sources(self.h, float(dt), ctypes.byref(src)), "afd_post_step_sources")

    def count_nonfinite(self):
        n = ctypes.c_int64()
        L.check(self.lib.afd_count_nonfinite(self.h, ctypes.byref(n)), "afd_count_nonfinite")
        return n.value

    def p2c(self, prim):
        prim = L.f64c(prim).reshape(-1, self.nb)
        out = np.empty_like(prim)
        L.check(self.lib.afd_p2c_batch(self.h, prim.shape[0], L.dptr(prim), L.dptr(out)), "afd_p2c_batch")
        return out

    def c2p(self, cons, with_iters=False):
        cons = L.f64c(cons).reshape(-1, self.nb)
        out = np.empty_like(cons)
        it = np.zeros(cons.shape[0], dtype=np.int32)
        L.check(self.lib.afd_c2p_batch(self.h, cons.shape[0], L.dptr(cons), L.dptr(out),
                                       it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "afd_c2p_batch")
        return (out, it) if with_iters else out

    # ---- diagnostics (utils/diagnostics.py)
    def diagnostics(self):
        d = L.Diag()
        L.check(self.lib.afd_diagnostics(self.h, ctypes.byref(d)), "afd_diagnostics")
        return {k: getattr(d, k) for k, _ in d._fields_}

    def centerline(self):
        """(Gamma, rho, p, vx) along x at the mid (j, k), interior cells: 4 x nx."""
        out = np.empty((4, self.params.nx))
        L.check(self.lib.afd_centerline(self.h, L.dptr(out), out.size), "afd_centerline")
        return out

    # ---- instrumentation
    def set_instrumentation(self, stats=True, timing=False):
        L.check(self.lib.afd_set_instrumentation(self.h, int(stats), int(timing)), "afd_set_instrumentation")

    def stats(self):
        st = L.Stats()
        L.check(self.lib.afd_get_stats(self.h, ctypes.byref(st)), "afd_get_stats")
        return st.as_dict()

    def c2p_flags(self):
        """Per-cell con2prim branch bits of the last stage of the latest step run with stats on
        (bit 0 hydro energy fallback, 1 bisection, 2 invalid recovery, 3 learned/previous start)."""
        n = int(np.prod(self.shape[1:]))
        out = np.zeros(n, dtype=np.uint8)
        L.check(self.lib.afd_get_c2p_flags(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n),
                "afd_get_c2p_flags")
        return out.reshape(self.shape[1:])

    def reset_stats(self):
        L.check(self.lib.afd_reset_stats(self.h), "afd_reset_stats")

    def synchronize(self):
        L.check(self.lib.afd_synchronize(self.h), "afd_synchronize")

    # ---- learned guess
    def set_mlp_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        L.check(self.lib.afd_mlp_set_weights(self.h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size),
                "afd_mlp_set_weights")

    def get_mlp_weights(self):
        n = ctypes.c_int32()
        L.check(self.lib.afd_mlp_param_count(ctypes.byref(n)), "afd_mlp_param_count")
        w = np.empty(n.value, dtype=np.float32)
        L.check(self.lib.afd_mlp_get_weights(self.h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size),
                "afd_mlp_get_weights")
        return w

    def mlp_dataset_from_prims(self, prim):
        """(features, targets, valid) of the con2prim pairs of primitive states (n x 9)."""
        prim = L.f64c(prim).reshape(-1, 9)
        n = prim.shape[0]
        X = np.empty((n, 5), dtype=np.float32)
        Y = np.empty((n, 2), dtype=np.float32)
        valid = np.empty(n, dtype=np.int32)
        L.check(self.lib.afd_mlp_dataset_from_prims(self.h, n, L.dptr(prim), _fptr(X), _fptr(Y),
                                                    valid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                "afd_mlp_dataset_from_prims")
        return X, Y, valid.astype(bool)

    def mlp_set_dataset(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 5)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        if X.shape[0] != Y.shape[0]:
            raise ValueError("X and Y hold different numbers of samples")
        L.check(self.lib.afd_mlp_set_dataset(self.h, X.shape[0], _fptr(X), _fptr(Y)), "afd_mlp_set_dataset")

    def mlp_train(self, steps, batch, lr):
        """`steps` Adam steps on consecutive minibatches; returns each minibatch's loss."""
        hist = np.zeros(max(int(steps), 1))
        L.check(self.lib.afd_mlp_train(self.h, int(steps), int(batch), float(lr), L.dptr(hist)), "afd_mlp_train")
        return hist[:int(steps)]

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 5)
        Y = np.empty((X.shape[0], 2), dtype=np.float32)
        L.check(self.lib.afd_mlp_forward(self.h, X.shape[0], _fptr(X), _fptr(Y)), "afd_mlp_forward")
        return Y

    def mlp_loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 5)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        n = ctypes.c_int32()
        L.check(self.lib.afd_mlp_param_count(ctypes.byref(n)), "afd_mlp_param_count")
        g = np.empty(n.value, dtype=np.float32)
        loss = np.zeros(1)
        L.check(self.lib.afd_mlp_loss_grad(self.h, X.shape[0], _fptr(X), _fptr(Y), _fptr(g), L.dptr(loss)),
                "afd_mlp_loss_grad")
        return float(loss[0]), g


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
