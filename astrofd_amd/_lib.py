"""ctypes binding of libafd.so (include/afd.h).

The library is the only compute path: importing this module on a machine where
libafd.so is missing raises immediately, and every call that returns an error
code raises AfdError with the library's message.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "libafd.so")

AFD_PHYS_HYDRO, AFD_PHYS_RMHD = 0, 1
AFD_PREC_FP64, AFD_PREC_MIXED = 0, 1
AFD_C2P_PLAIN, AFD_C2P_LEARNED, AFD_C2P_PREVIOUS = 0, 1, 2
AFD_BC_PERIODIC, AFD_BC_OUTFLOW = 0, 1
AFD_OK, AFD_ERR_INVALID, AFD_ERR_NOMEM, AFD_ERR_HIP, AFD_ERR_STATE = 0, -1, -2, -3, -4
AFD_FACE_PATH_DEFAULT = 3  # the face-kernel variant afd_params.face_path = 0 selects
AFD_NOALLOC_SCRATCH_BYTES = 16 << 20  # device staging of afd_c2p_batch_noalloc, per device


def c2p_noalloc_max(nb):
    """AFD_C2P_NOALLOC_MAX(nb): states one afd_c2p_batch_noalloc call can convert."""
    return AFD_NOALLOC_SCRATCH_BYTES // (2 * 8 * nb + 4)

_i32, _i64, _dbl, _f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_float
_D = ctypes.POINTER(ctypes.c_double)


class AfdError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("physics", _i32), ("nv", _i32), ("nx", _i32), ("ny", _i32), ("nz", _i32), ("ng", _i32),
        ("dx", _dbl), ("dy", _dbl), ("dz", _dbl),
        ("gamma", _dbl), ("p_max", _dbl), ("v_max", _dbl), ("glm_ch", _dbl), ("glm_cp", _dbl),
        ("limiter_id", _i32), ("recon_id", _i32), ("riemann_id", _i32), ("rk_order", _i32),
        ("n_passive", _i32), ("passive_offset", _i32),
        ("n_resist", _i32), ("resistivity", _dbl),
        ("precision", _i32), ("c2p_guess", _i32), ("bc_y", _i32), ("bc_z", _i32),
        ("two_temperature", _i32), ("thermo_offset", _i32), ("tei_tau", _dbl),
        ("device", _i32),
        ("gr_metric", _i32), ("gr_orthonormal", _i32), ("gr_x0", _i32), ("gr_mass", _dbl), ("gr_gamma", _dbl),
        ("face_path", _i32),
    ]


class Nozzle(ctypes.Structure):
    _fields_ = [
        ("y0", _dbl), ("z0", _dbl), ("jet_radius", _dbl), ("shear_thick", _dbl), ("profile", _i32),
        ("eta_rho", _dbl), ("rho_amb", _dbl), ("p_eq", _dbl), ("gamma_jet", _dbl),
        ("vx_amb", _dbl), ("vy_amb", _dbl), ("vz_amb", _dbl),
        ("perturb", _i32), ("turb_vamp", _dbl), ("turb_pamp", _dbl),
        ("n_tracers", _i32), ("tracer_offset", _i32),
        ("tracer_nozzle", _dbl * 8), ("tracer_amb", _dbl * 8),
        ("two_temperature", _i32), ("thermo_offset", _i32), ("te_nozzle", _dbl), ("ti_nozzle", _dbl),
        ("b_init", _i32), ("b0", _dbl), ("b_amb", _dbl * 3),
        ("chemistry", _i32), ("chem_offset", _i32), ("chem_nozzle", _dbl * 3),
    ]


class Chem(ctypes.Structure):
    _fields_ = [("offset", _i32), ("X", _dbl), ("Y", _dbl), ("tunit_k", _dbl), ("rate_scale", _dbl),
                ("energy_scale", _dbl), ("tmin_k", _dbl), ("gamma", _dbl)]


class Diss(ctypes.Structure):
    _fields_ = [("zeta", _dbl), ("eta", _dbl), ("kappa", _dbl), ("tau_bulk", _dbl), ("tau_shear", _dbl),
                ("tau_heat", _dbl), ("cap_frac", _dbl), ("p_max", _dbl), ("advect", _i32)]


class Sources(ctypes.Structure):
    _fields_ = [("dissipation", _i32), ("diss", Diss),
                ("cooling", _i32), ("cool", _dbl), ("heat", _dbl), ("p_cap", _dbl),
                ("chemistry", _i32), ("chem", Chem)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("hydro_iters", ctypes.c_uint64), ("newton_evals", ctypes.c_uint64),
        ("hydro_fallbacks", ctypes.c_uint64), ("bisections", ctypes.c_uint64), ("invalid", ctypes.c_uint64),
        ("learned_ok", ctypes.c_uint64), ("c2p_cells", ctypes.c_uint64),
        ("ms_face", _dbl * 3), ("ms_update", _dbl), ("ms_other", _dbl),
        ("launches_face", ctypes.c_uint64 * 3), ("launches_update", ctypes.c_uint64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("ms_face") and k != "launches_face"}
        d["ms_face"] = list(self.ms_face)
        d["launches_face"] = list(self.launches_face)
        return d


class Diag(ctypes.Structure):
    _fields_ = [("max_gamma", _dbl), ("inlet_flux", _dbl), ("divb_max", _dbl), ("divb_rms", _dbl),
                ("divb_mean", _dbl)]


_VP = ctypes.c_void_p
# name -> argtypes (all return int32)
SIGNATURES = {
    "afd_create": [ctypes.POINTER(Params), ctypes.POINTER(_VP)],
    "afd_destroy": [_VP],
    "afd_set_prim": [_VP, _D, _i64],
    "afd_get_prim": [_VP, _D, _i64],
    "afd_set_nozzle": [_VP, ctypes.POINTER(Nozzle)],
    "afd_nozzle_draw_count": [_VP, ctypes.POINTER(_i64)],
    "afd_apply_bcs": [_VP, _D, _i64],
    "afd_apply_periodic_yz": [_VP],
    "afd_apply_outflow_right_x": [_VP],
    "afd_apply_nozzle_left_x": [_VP, _D, _i64],
    "afd_max_char_speed": [_VP, ctypes.POINTER(_dbl)],
    "afd_step": [_VP, _dbl],
    "afd_compute_rhs": [_VP, _D, _i64],
    "afd_two_temperature": [_VP, _dbl],
    "afd_cooling_heating": [_VP, _dbl, _dbl, _dbl, _dbl],
    "afd_ion_chemistry": [_VP, _dbl, ctypes.POINTER(Chem)],
    "afd_post_step_sources": [_VP, _dbl, ctypes.POINTER(Sources)],
    "afd_causal_dissipation": [_VP, _dbl, ctypes.POINTER(Diss)],
    "afd_count_nonfinite": [_VP, ctypes.POINTER(_i64)],
    "afd_p2c_batch": [_VP, _i64, _D, _D],
    "afd_c2p_batch": [_VP, _i64, _D, _D, ctypes.POINTER(_i32)],
    "afd_c2p_batch_noalloc": [_VP, _i64, _D, _D, ctypes.POINTER(_i32)],
    "afd_diagnostics": [_VP, ctypes.POINTER(Diag)],
    "afd_centerline": [_VP, _D, _i64],
    "afd_set_instrumentation": [_VP, _i32, _i32],
    "afd_get_stats": [_VP, ctypes.POINTER(Stats)],
    "afd_reset_stats": [_VP],
    "afd_get_c2p_flags": [_VP, ctypes.POINTER(ctypes.c_uint8), _i64],
    "afd_synchronize": [_VP],
    "afd_mlp_param_count": [ctypes.POINTER(_i32)],
    "afd_mlp_set_weights": [_VP, ctypes.POINTER(_f32), _i32],
    "afd_mlp_get_weights": [_VP, ctypes.POINTER(_f32), _i32],
    "afd_mlp_dataset_from_prims": [_VP, _i64, _D, ctypes.POINTER(_f32), ctypes.POINTER(_f32), ctypes.POINTER(_i32)],
    "afd_mlp_set_dataset": [_VP, _i64, ctypes.POINTER(_f32), ctypes.POINTER(_f32)],
    "afd_mlp_train": [_VP, _i64, _i32, _f32, _D],
    "afd_mlp_forward": [_VP, _i64, ctypes.POINTER(_f32), ctypes.POINTER(_f32)],
    "afd_mlp_loss_grad": [_VP, _i64, ctypes.POINTER(_f32), ctypes.POINTER(_f32), ctypes.POINTER(_f32), _D],
}
STRING_FUNCS = ["afd_last_error", "afd_version"]

_lib = None


def load():
    """Load libafd.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfdError(f"{LIB_PATH} is missing: build it with `python -m astrofd_amd.build` "
                           "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i32
        for name in STRING_FUNCS:
            f = getattr(L, name)
            f.argtypes = []
            f.restype = ctypes.c_char_p
        _lib = L
    return _lib


def exported_symbols():
    return list(SIGNATURES) + STRING_FUNCS


def check(rc, what=""):
    if rc != 0:
        msg = load().afd_last_error().decode(errors="replace")
        raise AfdError(f"{what}: error {rc}: {msg}")


def dptr(a):
    return a.ctypes.data_as(_D)


def f64c(a):
    return np.ascontiguousarray(a, dtype=np.float64)
