"""The C-ABI library loads and exports every entry point include/afd.h declares (no GPU needed)."""
import ctypes
import os
import re

from astrofd_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "afd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(afd_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "afd_create" in names and "afd_step" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"libafd.so lacks {missing}"


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(L.exported_symbols())


def test_binding_constants_match_header():
    text = open(os.path.join(ROOT, "include", "afd.h")).read()
    for name in ("AFD_FACE_PATH_DEFAULT",):
        m = re.search(rf"^#define {name} (\d+)", text, flags=re.M)
        assert m and int(m.group(1)) == getattr(L, name), name


def test_version_and_param_count_without_device():
    lib = L.load()
    assert b"gfx950" in lib.afd_version()
    n = ctypes.c_int32()
    assert lib.afd_mlp_param_count(ctypes.byref(n)) == 0
    assert n.value == 4674


def test_invalid_params_rejected_before_touching_a_device():
    lib = L.load()
    p = L.Params()
    p.physics = 7
    h = ctypes.c_void_p()
    rc = lib.afd_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"physics" in lib.afd_last_error()


def _params(**kw):
    p = L.Params(physics=L.AFD_PHYS_HYDRO, nv=5, nx=8, ny=8, nz=8, ng=3, dx=0.1, dy=0.1, dz=0.1, gamma=5 / 3,
                 p_max=2.0, v_max=0.9999, rk_order=3, passive_offset=5, precision=L.AFD_PREC_FP64)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_parameter_validation_without_a_device():
    """Every check afd_create makes before it selects a device, with the message it leaves."""
    lib = L.load()
    cases = [
        (dict(ng=1), b"NG"),
        (dict(nx=1), b"grid too small"),
        (dict(nv=7, n_passive=1), b"nv must equal passive_offset + n_passive"),
        (dict(physics=L.AFD_PHYS_RMHD, nv=10, n_passive=1, passive_offset=10), b"RMHD passives"),
        (dict(recon_id=3), b"recon_id"),
        (dict(limiter_id=5), b"limiter_id"),
        (dict(riemann_id=2), b"riemann_id"),  # SRHD: hlle / hllc only
        (dict(rk_order=4), b"rk_order"),
        (dict(precision=9), b"precision"),
        (dict(c2p_guess=L.AFD_C2P_LEARNED), b"RMHD only"),
        (dict(two_temperature=1, thermo_offset=5, nv=6, n_passive=1), b"thermo_offset"),
        (dict(dx=0.0), b"dx"),
        (dict(n_resist=3), b"n_resist"),
        (dict(gr_metric=4), b"gr_metric"),
        (dict(gr_metric=1, gr_gamma=1.0), b"gr_gamma"),
        (dict(nv=40, n_passive=35), b"nv > 32"),
        (dict(nx=1285, ny=1285, nz=1285), b"grid too large"),  # 1291^3 >= 2^31 cells with ghosts
        (dict(face_path=5), b"face_path"),
    ]
    for kw, msg in cases:
        h = ctypes.c_void_p()
        rc = lib.afd_create(ctypes.byref(_params(**kw)), ctypes.byref(h))
        assert rc == L.AFD_ERR_INVALID and msg in lib.afd_last_error(), (kw, lib.afd_last_error())
        assert not h.value
    assert lib.afd_create(None, None) == L.AFD_ERR_INVALID


def test_null_context_calls_fail_cleanly():
    """Entry points given a null context return AFD_ERR_INVALID instead of dereferencing it."""
    lib = L.load()
    assert lib.afd_step(None, ctypes.c_double(0.1)) == L.AFD_ERR_INVALID
    assert lib.afd_two_temperature(None, ctypes.c_double(0.1)) == L.AFD_ERR_INVALID
    assert lib.afd_cooling_heating(None, 0.1, 1.0, 0.0, 1.0) == L.AFD_ERR_INVALID
    assert lib.afd_ion_chemistry(None, 0.1, None) == L.AFD_ERR_INVALID
    assert lib.afd_causal_dissipation(None, 0.1, None) == L.AFD_ERR_INVALID
    assert lib.afd_post_step_sources(None, 0.1, None) == L.AFD_ERR_INVALID
