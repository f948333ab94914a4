#!/usr/bin/env python3
"""Write the golden input/output vectors for the SRHD/RMHD finite-volume step.

This script runs ONLY in the development container, where the AstroFD reference
sources are mounted read-only at /root/reference.  It imports the reference's
pure-Python modules with a stand-in ``numba`` module whose ``njit`` returns the
function unchanged (numba itself is not installed), calls the reference
functions that no closed form or invariant pins on small seeded inputs, and
stores inputs and outputs as compressed ``.npz`` fixtures next to this file.

Nothing under ``tests/`` imports this script; the tests read only the fixtures.

Usage:
    python tests/golden/make_golden.py [--only NAME ...] [--driver]

``--driver`` additionally writes ``driver_srhd_32.npz`` (the 32^3, 20-step
SRHD MUSCL-minmod/HLLE/SSPRK2 jet with +-1% density noise, periodic y/z as in
the reference), which takes about ten minutes of pure-Python time.
"""
import argparse
import math
import os
import sys
import time
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _install_numba_stub():
    nb = types.ModuleType("numba")

    def njit(*args, **kwargs):
        if len(args) == 1 and callable(args[0]) and not kwargs:
            return args[0]

        def deco(f):
            return f
        return deco

    nb.njit = njit
    nb.prange = range
    sys.modules["numba"] = nb


_install_numba_stub()
sys.path.insert(0, REF)
from core import srhd_core, rmhd_core, boundary, nozzle, source_terms, chemistry, dissipation  # noqa: E402
from core import grhd_core, grmhd_core  # noqa: E402


def _save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"[golden] wrote {name}.npz ({os.path.getsize(path)/1024:.1f} KiB)", flush=True)


# ---------------------------------------------------------------------------
# Seeded primitive states
# ---------------------------------------------------------------------------
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from tests.golden_states import rmhd_state, smooth_state, source_state, srhd_state  # noqa: E402


SRHD_BASE = dict(GAMMA=5.0 / 3.0, P_MAX=2.0, V_MAX=0.9999, N_PASSIVE=0, PASSIVE_OFFSET=5)
RMHD_BASE = dict(GAMMA=4.0 / 3.0, P_MAX=2.0, V_MAX=0.9999, GLM_CH=1.0, GLM_CP=0.1,
                 N_PASSIVE=0, PASSIVE_OFFSET=9, RESISTIVE_ENABLED=False, RESISTIVITY=0.0)


def srhd_configure(**kw):
    p = dict(SRHD_BASE)
    p.update(kw)
    srhd_core.configure(p)
    return p


def rmhd_configure(**kw):
    p = dict(RMHD_BASE)
    p.update(kw)
    rmhd_core.configure(p)
    return p


def _param_record(p):
    keys = sorted(p.keys())
    return np.array([f"{k}={p[k]}" for k in keys])


# ---------------------------------------------------------------------------
# Pointwise functions
# ---------------------------------------------------------------------------
def gen_pointwise():
    rng = np.random.default_rng(11)
    n = 400
    # SRHD prim -> cons -> prim (c2p fixed point with fallback, Q4)
    srhd_configure(P_MAX=2.0, V_MAX=0.9999)
    rho = 10 ** rng.uniform(-3, 1, n)
    vmag = np.concatenate([np.zeros(20), 10 ** rng.uniform(-9, -5, 40), rng.uniform(0, 0.9999, n - 60)])
    th = np.arccos(rng.uniform(-1, 1, n)); ph = rng.uniform(0, 2 * np.pi, n)
    vx = vmag * np.sin(th) * np.cos(ph); vy = vmag * np.sin(th) * np.sin(ph); vz = vmag * np.cos(th)
    p = 10 ** rng.uniform(-4, 0.5, n)          # a few above P_MAX=2
    prim = np.stack([rho, vx, vy, vz, p], 1)
    cons = np.array([srhd_core.prim_to_cons(*prim[i]) for i in range(n)])
    # perturbed conserved states (off the p2c manifold) exercise the clamps
    cons2 = cons * (1.0 + 0.3 * rng.uniform(-1, 1, cons.shape))
    cons_all = np.concatenate([cons, cons2], 0)
    c2p = np.array([srhd_core.cons_to_prim(*cons_all[i]) for i in range(len(cons_all))])
    cs = np.array([srhd_core.sound_speed(prim[i, 0], prim[i, 4]) for i in range(n)])
    eig = np.array([srhd_core.eig_speeds(prim[i, 1], cs[i]) for i in range(n)])
    fl = np.array([np.concatenate([srhd_core.flux_x(*prim[i]), srhd_core.flux_y(*prim[i]),
                                   srhd_core.flux_z(*prim[i])]) for i in range(n)])
    # limiters / PPM / WENO on random 5-point stencils (with some ties and zeros)
    q = rng.normal(size=(n, 5))
    q[:40, 1] = q[:40, 2]
    q[40:60] = 0.0
    lim = {}
    for name, lid in (("mc", 0), ("minmod", 1), ("vanleer", 2)):
        srhd_core.LIMITER_ID = lid
        lim[name] = np.array([srhd_core.mc_lim(q[i, 0], q[i, 1]) for i in range(n)])
    srhd_core.LIMITER_ID = 0
    ppm = np.array([srhd_core.ppm_reconstruct(*q[i]) for i in range(n)])
    wl = np.array([srhd_core.weno5_left(*q[i]) for i in range(n)])
    wr = np.array([srhd_core.weno5_right(*q[i]) for i in range(n)])
    # SRHD Riemann solvers on random left/right states
    iL = rng.integers(0, n, n); iR = rng.integers(0, n, n)
    hll = []; hllc = []
    for m in range(n):
        a = prim[iL[m]]; b = prim[iR[m]]
        a = np.array(srhd_core.floor_prim(*a)); b = np.array(srhd_core.floor_prim(*b))
        UL = np.array(srhd_core.prim_to_cons(*a)); UR = np.array(srhd_core.prim_to_cons(*b))
        row_e = []; row_c = []
        for d, fn, hc in ((1, srhd_core.flux_x, srhd_core.hllc_flux_x), (2, srhd_core.flux_y, srhd_core.hllc_flux_y),
                          (3, srhd_core.flux_z, srhd_core.hllc_flux_z)):
            FL = fn(*a); FR = fn(*b)
            csL = srhd_core.sound_speed(a[0], a[4]); csR = srhd_core.sound_speed(b[0], b[4])
            lmL, lpL = srhd_core.eig_speeds(a[d], csL); lmR, lpR = srhd_core.eig_speeds(b[d], csR)
            sL = min(lmL, lmR); sR = max(lpL, lpR)
            row_e.append(srhd_core.hlle(UL, UR, FL, FR, sL, sR))
            row_c.append(hc(UL, UR, FL, FR, sL, sR, a[4], b[4], a[d], b[d]))
        hll.append(np.concatenate(row_e)); hllc.append(np.concatenate(row_c))
    _save("srhd_pointwise", prim=prim, cons=cons, cons_all=cons_all, c2p=c2p, cs=cs, eig=eig, flux=fl,
          q=q, lim_mc=lim["mc"], lim_minmod=lim["minmod"], lim_vanleer=lim["vanleer"], ppm=ppm,
          weno_l=wl, weno_r=wr, iL=iL, iR=iR, hlle=np.array(hll), hllc=np.array(hllc),
          params=_param_record(SRHD_BASE))

    # RMHD pointwise: p2c, hydro guess, f(Z), c2p (FD-Newton, bisection, invalid fallback), fast speed
    p = rmhd_configure(P_MAX=1.0e4, V_MAX=0.9999)
    m = 500
    rho = 10 ** rng.uniform(-4, 2, m)
    W = rng.uniform(1.0, 50.0, m)
    W[:30] = 1.0
    vmag = np.sqrt(1.0 - 1.0 / W ** 2)
    th = np.arccos(rng.uniform(-1, 1, m)); ph = rng.uniform(0, 2 * np.pi, m)
    vx = vmag * np.sin(th) * np.cos(ph); vy = vmag * np.sin(th) * np.sin(ph); vz = vmag * np.cos(th)
    pp = rho * 10 ** rng.uniform(-4, 2, m)
    sig = 10 ** rng.uniform(-3, 1, m)
    th2 = np.arccos(rng.uniform(-1, 1, m)); ph2 = rng.uniform(0, 2 * np.pi, m)
    bx = np.sin(th2) * np.cos(ph2); by = np.sin(th2) * np.sin(ph2); bz = np.cos(th2)
    h = 1.0 + 4.0 * pp / rho
    vb = vx * bx + vy * by + vz * bz
    Bm = np.sqrt(sig * rho * h / (1.0 / W ** 2 + vb ** 2))
    psi = 1e-3 * rng.uniform(-1, 1, m)
    primm = np.stack([rho, vx, vy, vz, pp, Bm * bx, Bm * by, Bm * bz, psi], 1)
    consm = np.array([rmhd_core.prim_to_cons_rmhd(*primm[i], rmhd_core.GAMMA) for i in range(m)])
    t0 = time.time()
    c2pm = np.array([rmhd_core.cons_to_prim_rmhd(*consm[i]) for i in range(m)])
    # states off the p2c manifold (negative tau, huge S) reach bisection and the hydro fallback
    consb = consm[:100].copy()
    consb[:, 1:4] *= rng.uniform(0.5, 3.0, (100, 1))
    consb[:50, 4] *= rng.uniform(-2.0, 0.5, 50)
    c2pb = np.array([rmhd_core.cons_to_prim_rmhd(*consb[i]) for i in range(100)])
    hyd = np.array([rmhd_core._cons_to_prim_hydro(*consm[i, :5], rmhd_core.GAMMA) for i in range(m)])
    Z = 10 ** rng.uniform(-3, 4, m)
    B2 = (consm[:, 5:8] ** 2).sum(1); SB = (consm[:, 1:4] * consm[:, 5:8]).sum(1); S2 = (consm[:, 1:4] ** 2).sum(1)
    fz = np.array([rmhd_core.rmhd_f_of_Z(Z[i], consm[i, 0], consm[i, 4], B2[i], SB[i], S2[i]) for i in range(m)])
    cf = np.array([rmhd_core.rmhd_fast_speed(primm[i, 0], primm[i, 4], *primm[i, 1:4], *primm[i, 5:8])
                   for i in range(m)])
    print(f"[golden] rmhd c2p pointwise {time.time()-t0:.1f}s", flush=True)
    _save("rmhd_pointwise", prim=primm, cons=consm, c2p=c2pm, cons_bad=consb, c2p_bad=c2pb, hydro=hyd,
          Z=Z, fZ=fz, cf=cf, params=_param_record(p))

    # RMHD Riemann solvers (HLLE, hlld hybrid, hlld_full) on floored random L/R states
    rmhd_configure(P_MAX=2.0, V_MAX=0.9999)
    st = rmhd_state((4, 5, 6), 77)
    flat = st.reshape(9, -1).T
    k = flat.shape[0]
    iL = rng.integers(0, k, 300); iR = rng.integers(0, k, 300)
    iR[:10] = iL[:10]  # identical states
    out = {0: [], 1: [], 2: []}
    for mm in range(300):
        a = np.array(rmhd_core.floor_prim_rmhd(*flat[iL[mm]]))
        b = np.array(rmhd_core.floor_prim_rmhd(*flat[iR[mm]]))
        if mm % 7 == 0:
            b[5] = -a[5]  # Bn averages to ~0 -> HLLE fallback
        UL = np.array(rmhd_core.prim_to_cons_rmhd(*a, rmhd_core.GAMMA))
        UR = np.array(rmhd_core.prim_to_cons_rmhd(*b, rmhd_core.GAMMA))
        for rid in (0, 1, 2):
            rmhd_core.RIEMANN_ID = rid
            row = []
            for d, fn, rf in ((1, rmhd_core.flux_rmhd_x, rmhd_core.riemann_flux_rmhd_x),
                              (2, rmhd_core.flux_rmhd_y, rmhd_core.riemann_flux_rmhd_y),
                              (3, rmhd_core.flux_rmhd_z, rmhd_core.riemann_flux_rmhd_z)):
                FL = fn(a); FR = fn(b)
                cfL = rmhd_core.rmhd_fast_speed(a[0], a[4], a[1], a[2], a[3], a[5], a[6], a[7])
                cfR = rmhd_core.rmhd_fast_speed(b[0], b[4], b[1], b[2], b[3], b[5], b[6], b[7])
                lmL, lpL = rmhd_core.rmhd_eig_speeds(a[d], cfL); lmR, lpR = rmhd_core.rmhd_eig_speeds(b[d], cfR)
                sL = min(lmL, lmR); sR = max(lpL, lpR)
                if mm % 5 == 1:
                    sL, sR = -1.0, 1.0  # the carried (-1, 1) seed of the PPM/WENO x faces
                row.append(rf(UL, UR, FL, FR, sL, sR, a[4], b[4], a[d], b[d]))
            out[rid].append(np.concatenate(row))
    rmhd_core.RIEMANN_ID = 0
    _save("rmhd_riemann", states=flat, iL=iL, iR=iR, hlle=np.array(out[0]), hlld=np.array(out[1]),
          hlld_full=np.array(out[2]))


# ---------------------------------------------------------------------------
# Whole-array RHS and SSPRK steps
# ---------------------------------------------------------------------------
SRHD_RHS_CASES = [
    ("muscl", "mc", "hlle"), ("muscl", "minmod", "hlle"), ("muscl", "vanleer", "hlle"),
    ("muscl", "mc", "hllc"), ("muscl", "vanleer", "hllc"), ("muscl", "minmod", "hllc"),
    ("ppm", "mc", "hlle"), ("ppm", "mc", "hllc"), ("weno", "mc", "hlle"), ("weno", "mc", "hllc"),
]
RMHD_RHS_CASES = [
    ("muscl", "mc", "hlle"), ("muscl", "minmod", "hlld"), ("muscl", "vanleer", "hlld_full"),
    ("muscl", "mc", "hlld"), ("muscl", "mc", "hlld_full"),
    ("ppm", "mc", "hlle"), ("ppm", "mc", "hlld"), ("ppm", "mc", "hlld_full"),
    ("weno", "mc", "hlle"), ("weno", "mc", "hlld"), ("weno", "mc", "hlld_full"),
]
RECON_ID = {"muscl": 0, "ppm": 1, "weno": 2}


def srhd_rhs(pr, recon, dx, dy, dz):
    n = pr.shape[1:]
    if recon == "ppm":
        return srhd_core.compute_rhs_ppm(pr, *n, dx, dy, dz)
    if recon == "weno":
        return srhd_core.compute_rhs_weno(pr, *n, dx, dy, dz)
    return srhd_core.compute_rhs_muscl(pr, *n, dx, dy, dz)


def rmhd_rhs(pr, recon, dx, dy, dz):
    n = pr.shape[1:]
    if recon == "ppm":
        return rmhd_core.compute_rhs_ppm(pr, *n, dx, dy, dz)
    if recon == "weno":
        return rmhd_core.compute_rhs_weno(pr, *n, dx, dy, dz)
    return rmhd_core.compute_rhs_rmhd(pr, *n, dx, dy, dz)


def gen_rhs():
    shape = (11, 12, 13)     # 5-7 interior cells with NG = 3, all axes distinct
    dx, dy, dz = 1.0 / 5, 1.0 / 6, 1.0 / 7
    arrs = {}
    for c, (recon, lim, riem) in enumerate(SRHD_RHS_CASES):
        srhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem)
        pr = srhd_state(shape, 100 + c)
        t0 = time.time()
        rhs = srhd_rhs(pr, recon, dx, dy, dz)
        arrs[f"pr_{c}"] = pr; arrs[f"rhs_{c}"] = rhs
        arrs[f"amax_{c}"] = np.array(srhd_core.max_char_speed(pr, *shape))
        print(f"[golden] srhd rhs {recon}/{lim}/{riem} {time.time()-t0:.1f}s", flush=True)
    # passive scalars (Q15): MUSCL x triple subtraction with wrapped i-3 index, PPM y/z only
    for c, recon in ((len(SRHD_RHS_CASES), "muscl"), (len(SRHD_RHS_CASES) + 1, "ppm")):
        srhd_configure(RECON=recon, LIMITER="vanleer", RIEMANN="hllc", N_PASSIVE=2, PASSIVE_OFFSET=5)
        pr = srhd_state(shape, 100 + c, npass=2)
        rhs = srhd_rhs(pr, recon, dx, dy, dz)
        arrs[f"pr_{c}"] = pr; arrs[f"rhs_{c}"] = rhs
        arrs[f"amax_{c}"] = np.array(srhd_core.max_char_speed(pr, *shape))
        print(f"[golden] srhd rhs passive {recon}", flush=True)
    cases = SRHD_RHS_CASES + [("muscl", "vanleer", "hllc"), ("ppm", "vanleer", "hllc")]
    _save("srhd_rhs", cases=np.array(["/".join(c) for c in cases]), npass=np.array([0] * 10 + [2, 2]),
          dxyz=np.array([dx, dy, dz]), **arrs)

    arrs = {}
    for c, (recon, lim, riem) in enumerate(RMHD_RHS_CASES):
        rmhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem)
        pr = rmhd_state(shape, 200 + c)
        t0 = time.time()
        rhs = rmhd_rhs(pr, recon, dx, dy, dz)
        arrs[f"pr_{c}"] = pr; arrs[f"rhs_{c}"] = rhs
        print(f"[golden] rmhd rhs {recon}/{lim}/{riem} {time.time()-t0:.1f}s", flush=True)
    # resistivity: once in MUSCL, twice in PPM, never in WENO (Q9)
    extra = [("muscl", "mc", "hlle"), ("ppm", "mc", "hlld_full"), ("weno", "mc", "hlld_full")]
    for e, (recon, lim, riem) in enumerate(extra):
        c = len(RMHD_RHS_CASES) + e
        rmhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, RESISTIVE_ENABLED=True, RESISTIVITY=1e-3)
        pr = rmhd_state(shape, 200 + c)
        rhs = rmhd_rhs(pr, recon, dx, dy, dz)
        arrs[f"pr_{c}"] = pr; arrs[f"rhs_{c}"] = rhs
        print(f"[golden] rmhd rhs resistive {recon}", flush=True)
    cases = RMHD_RHS_CASES + extra
    _save("rmhd_rhs", cases=np.array(["/".join(c) for c in cases]),
          resistive=np.array([0] * len(RMHD_RHS_CASES) + [1] * len(extra)), dxyz=np.array([dx, dy, dz]), **arrs)


class WrapPastEnd(np.ndarray):
    """A view whose scalar reads past the end of an axis wrap to its start.  The reference's RMHD
    MUSCL passive stencils read pr[idx, i+3] at i = nx-3 (and j+3, k+3 likewise), which is an
    IndexError in pure Python and a read of the next row in Numba; the GPU library wraps such a
    read.  Negative indices keep Python's own wrap."""

    def __getitem__(self, key):
        if isinstance(key, tuple) and all(isinstance(k, (int, np.integer)) for k in key):
            key = tuple(int(k) % n if k >= n else k for k, n in zip(key, self.shape))
        return np.ndarray.__getitem__(self, key)


RMHD_PASSIVE_RHS_CASES = [("muscl", "mc", "hlle"), ("muscl", "vanleer", "hlld"), ("muscl", "minmod", "hlld_full")]


def gen_rhs_passive():
    """RMHD MUSCL right-hand sides with two passive scalars (core/rmhd_core.py:832-940 and the y/z
    counterparts: every variant's fluxes, right faces with the left face's speeds)."""
    shape = (11, 12, 13)
    dx, dy, dz = 1.0 / 5, 1.0 / 6, 1.0 / 7
    arrs = {}
    for c, (recon, lim, riem) in enumerate(RMHD_PASSIVE_RHS_CASES):
        rmhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, N_PASSIVE=2, PASSIVE_OFFSET=9)
        pr = smooth_state(shape, 600 + c, "rmhd", vamp=0.2, npass=2)
        t0 = time.time()
        rhs = rmhd_core.compute_rhs_rmhd(pr.view(WrapPastEnd), *shape, dx, dy, dz)
        arrs[f"pr_{c}"] = pr; arrs[f"rhs_{c}"] = np.asarray(rhs)
        print(f"[golden] rmhd rhs passive {recon}/{lim}/{riem} {time.time()-t0:.1f}s", flush=True)
    _save("rmhd_rhs_passive", cases=np.array(["/".join(c) for c in RMHD_PASSIVE_RHS_CASES]),
          dxyz=np.array([dx, dy, dz]), **arrs)


SRHD_STEP_CASES = [  # recon, limiter, riemann, rk, npass
    ("muscl", "minmod", "hlle", 2, 0), ("muscl", "vanleer", "hllc", 3, 0),
    ("ppm", "mc", "hlle", 2, 0), ("weno", "mc", "hllc", 3, 0), ("muscl", "mc", "hlle", 2, 1),
]
RMHD_STEP_CASES = [  # recon, limiter, riemann, rk, npass, resistive
    ("weno", "mc", "hlld_full", 3, 0, False), ("muscl", "minmod", "hlld", 2, 0, False),
    ("ppm", "mc", "hlle", 3, 0, False), ("weno", "mc", "hlld_full", 3, 2, True),
    ("muscl", "vanleer", "hlld_full", 2, 0, True),
]


def gen_steps():
    shape = (11, 12, 13)
    dx, dy, dz = 1.0 / 5, 1.0 / 6, 1.0 / 7
    dt = 0.02
    arrs = {}
    for c, (recon, lim, riem, rk, npass) in enumerate(SRHD_STEP_CASES):
        srhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, N_PASSIVE=npass, PASSIVE_OFFSET=5)
        pr = srhd_state(shape, 300 + c, npass=npass)
        t0 = time.time()
        out = (srhd_core.step_ssprk3 if rk == 3 else srhd_core.step_ssprk2)(pr, dx, dy, dz, dt)
        arrs[f"pr_{c}"] = pr; arrs[f"out_{c}"] = out
        print(f"[golden] srhd step {recon}/{lim}/{riem}/rk{rk}/np{npass} {time.time()-t0:.1f}s", flush=True)
    _save("srhd_step", cases=np.array(["/".join(map(str, c)) for c in SRHD_STEP_CASES]),
          dxyz=np.array([dx, dy, dz]), dt=np.array(dt), **arrs)

    arrs = {}
    for c, (recon, lim, riem, rk, npass, res) in enumerate(RMHD_STEP_CASES):
        rmhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, N_PASSIVE=npass, PASSIVE_OFFSET=9,
                       RESISTIVE_ENABLED=res, RESISTIVITY=1e-3 if res else 0.0)
        # smooth states: rough ones make the reference RMHD step diverge within one step
        pr = smooth_state(shape, 400 + c, "rmhd", vamp=0.2, npass=npass)
        t0 = time.time()
        out = (rmhd_core.step_ssprk3 if rk == 3 else rmhd_core.step_ssprk2)(pr, dx, dy, dz, dt)
        print(f"[golden]   out rho [{out[0].min():.3e}, {out[0].max():.3e}] p max {out[4].max():.3e} "
              f"|B| max {np.abs(out[5:8]).max():.3e}", flush=True)
        arrs[f"pr_{c}"] = pr; arrs[f"out_{c}"] = out
        print(f"[golden] rmhd step {recon}/{lim}/{riem}/rk{rk}/np{npass}/res{res} {time.time()-t0:.1f}s",
              flush=True)
    _save("rmhd_step", cases=np.array(["/".join(map(str, c)) for c in RMHD_STEP_CASES]),
          dxyz=np.array([dx, dy, dz]), dt=np.array(dt), **arrs)


# ---------------------------------------------------------------------------
# Boundary / nozzle fills and the two-temperature source
# ---------------------------------------------------------------------------
NOZZLE_CASES = [
    # physics, profile, shear, perturb, ntracers, two_temp, b_init
    ("hydro", "top_hat", 0.02, "random", 1, False, "none"),
    ("hydro", "top_hat", 0.0, "none", 0, False, "none"),
    ("hydro", "taper", 0.0, "random", 2, True, "none"),
    ("rmhd", "parabolic", 0.0, "none", 0, False, "toroidal"),
    ("rmhd", "top_hat", 0.05, "random", 0, True, "poloidal"),
]


def gen_bc():
    ng = 3
    NX, NY, NZ = 4, 7, 6
    shape = (NX + 2 * ng, NY + 2 * ng, NZ + 2 * ng)
    dx, dy, dz = 1.0 / NX, 1.0 / NY, 1.0 / NZ
    arrs = {}
    for c, (phys, prof, shear, pert, ntr, tt, binit) in enumerate(NOZZLE_CASES):
        nv_base = 9 if phys == "rmhd" else 5
        nthermo = 2 if tt else 0
        npass = ntr + nthermo
        cfg = dict(NG=ng, JET_RADIUS=0.3, SHEAR_THICK=shear, NOZZLE_PROFILE=prof, GAMMA_JET=7.0888121,
                   ETA_RHO=0.01, P_EQ=0.01, RHO_AMB=1.0, VX_AMB=0.01, VY_AMB=-0.02, VZ_AMB=0.03,
                   NOZZLE_PERTURB=pert, TURB_VAMP=0.02, TURB_PAMP=0.1, N_TRACERS=ntr, TRACER_OFFSET=nv_base,
                   TRACER_NOZZLE_VALUES=[1.0, 0.5][:ntr], TRACER_AMB_VALUES=[0.0, 0.25][:ntr],
                   TWO_TEMPERATURE=tt, THERMO_OFFSET=nv_base + ntr, TE_NOZZLE=0.7, TI_NOZZLE=0.9,
                   PHYSICS=phys, B_INIT=binit, B0=0.5)
        pr = (rmhd_state if phys == "rmhd" else srhd_state)(shape, 500 + c, npass=npass)
        before = pr.copy()
        boundary.apply_periodic_yz(pr, ng)
        after_per = pr.copy()
        boundary.apply_outflow_right_x(pr, ng)
        after_out = pr.copy()
        rng = np.random.default_rng(1234)
        nozzle.apply_nozzle_left_x(pr, dx, dy, dz, NY, NZ, 0.45, 0.55, rng, cfg)
        nozzle.apply_nozzle_left_x(pr, dx, dy, dz, NY, NZ, 0.45, 0.55, rng, cfg)  # second call: next draws
        after_noz = pr.copy()
        tt_cfg = dict(TWO_TEMPERATURE=tt, N_TRACERS=ntr, THERMO_OFFSET=nv_base + ntr, TEI_TAU=0.1, NG=ng)
        source_terms.apply_two_temperature(pr, 0.01, tt_cfg)
        # the periodic and outflow copies are pinned by the final ghost contents (corners included)
        del after_per, after_out
        arrs[f"before_{c}"] = before; arrs[f"nozzle_{c}"] = after_noz; arrs[f"twotemp_{c}"] = pr.copy()
    weights = {}
    r = np.linspace(0.0, 0.6, 61)
    for prof in ("top_hat", "taper", "parabolic"):
        for sh in (0.0, 0.02):
            weights[f"w_{prof}_{sh}"] = np.array([nozzle.nozzle_weight(x, 0.3, sh, prof) for x in r])
    _save("bc_nozzle", cases=np.array(["/".join(map(str, c)) for c in NOZZLE_CASES]), r=r,
          grid=np.array([NX, NY, NZ, ng]), **arrs, **weights)


# ---------------------------------------------------------------------------
# Post-step sources (core/source_terms.py, core/chemistry.py) and the nozzle's
# chemistry fill (core/nozzle.py:77-82)
# ---------------------------------------------------------------------------
COOLING_CASES = [
    # dt, settings (P_MAX absent: the reference's default 1.0)
    (0.05, dict(COOLING_ENABLED=True, COOLING_LAMBDA=5.0, HEATING_RATE=0.0, P_MAX=2.0)),
    (0.5, dict(COOLING_ENABLED=True, COOLING_LAMBDA=0.0, HEATING_RATE=0.3)),
    (0.1, dict(COOLING_ENABLED=True, COOLING_LAMBDA=50.0, HEATING_RATE=0.01, P_MAX=2.0)),
    (0.1, dict(COOLING_ENABLED=True, COOLING_LAMBDA=0.0, HEATING_RATE=0.0)),
    (0.1, dict(COOLING_ENABLED=False, COOLING_LAMBDA=5.0, HEATING_RATE=0.2)),
]
CHEM_CASES = [
    (0.01, dict()),
    (0.05, dict(CHEM_TUNIT_K=1.0e6, CHEM_RATE_SCALE=1.0e9, CHEM_ENERGY_SCALE=1.0e-7, CHEM_TMIN_K=5.0e3,
                GAMMA=4.0 / 3.0)),
    (0.1, dict(CHEM_TUNIT_K=1.0e7, CHEM_RATE_SCALE=1.0e12, CHEM_ENERGY_SCALE=1.0e-4)),
    (1e-3, dict(CHEM_X=0.7, CHEM_Y=0.3, CHEM_TUNIT_K=1.0e5, CHEM_RATE_SCALE=1.0e10, CHEM_ENERGY_SCALE=1.0e-6)),
    # strong coupling: the pressure floor (1e-20) and the clamps of xHII and xHeIII at 1 act
    (0.05, dict(CHEM_TUNIT_K=1.0e8, CHEM_RATE_SCALE=1.0e10, CHEM_ENERGY_SCALE=10.0)),
]


def gen_sources():
    ng = 2
    shape = (4 + 2 * ng, 5 + 2 * ng, 6 + 2 * ng)
    arrs = {}
    # cooling/heating: rho in (0.01, 2), p in (1e-6, 1.5)
    for c, (dt, cfg) in enumerate(COOLING_CASES):
        rng = np.random.default_rng(700 + c)
        pr = np.zeros((5,) + shape)
        pr[0] = rng.uniform(0.01, 2.0, shape)
        pr[4] = 10.0 ** rng.uniform(-6, np.log10(1.5), shape)
        arrs[f"cool_in_{c}"] = pr.copy()
        source_terms.apply_cooling_heating(pr, dt, dict(cfg, NG=ng))
        arrs[f"cool_out_{c}"] = pr[4]  # the only variable it writes
    # ion chemistry on hydro (nv = 8) and RMHD (nv = 12) layouts
    for c, (dt, cfg) in enumerate(CHEM_CASES):
        nvb = 9 if c % 2 else 5
        pr = source_state(shape, 720 + c, nvb + 3, chem_off=nvb)
        full = dict(cfg, CHEMISTRY_ENABLED=True, N_CHEM=3, CHEM_OFFSET=nvb, NG=ng)
        arrs[f"chem_in_{c}"] = pr.copy()
        chemistry.apply_ion_chemistry(pr, dt, full)
        arrs[f"chem_out_{c}"] = pr[[4, nvb, nvb + 1, nvb + 2]]  # p and the fractions
    # the driver's post-step order (solvers/srhd3d_mpi_muscl.py:324-326): cooling, two-temperature,
    # chemistry, on one hydro block with a tracer, Te/Ti and the ion fractions
    cfg = dict(NG=ng, COOLING_ENABLED=True, COOLING_LAMBDA=2.0, HEATING_RATE=0.05, P_MAX=50.0,
               TWO_TEMPERATURE=True, N_TRACERS=1, THERMO_OFFSET=6, TEI_TAU=0.2,
               CHEMISTRY_ENABLED=True, N_CHEM=3, CHEM_OFFSET=8, CHEM_TUNIT_K=1.0e6, CHEM_RATE_SCALE=1.0e9,
               CHEM_ENERGY_SCALE=1.0e-8)
    pr = source_state(shape, 740, 11, chem_off=8, bad=False)
    arrs["chain_in"] = pr.copy()
    source_terms.apply_cooling_heating(pr, 0.02, cfg)
    source_terms.apply_two_temperature(pr, 0.02, cfg)
    chemistry.apply_ion_chemistry(pr, 0.02, cfg)
    arrs["chain_out"] = pr[[4, 6, 7, 8, 9, 10]]  # p, Te, Ti and the fractions
    # nozzle fill with chemistry fractions; the 16-variable layouts also take the pr[5:15] = 0 branch
    NX, NY, NZ = 2, 6, 5
    nshape = (NX + 2 * ng, NY + 2 * ng, NZ + 2 * ng)
    noz_cases = [
        # physics, n_tracers, two_temperature, chem_offset, nv, b_init
        ("hydro", 1, True, 8, 11, "none"),
        ("rmhd", 2, True, 13, 16, "poloidal"),
        ("hydro", 3, True, 13, 16, "none"),
    ]
    for c, (phys, ntr, tt, coff, nv, binit) in enumerate(noz_cases):
        nvb = 9 if phys == "rmhd" else 5
        cfg = dict(NG=ng, JET_RADIUS=0.3, SHEAR_THICK=0.02, NOZZLE_PROFILE="top_hat", GAMMA_JET=7.0888121,
                   ETA_RHO=0.01, P_EQ=0.01, RHO_AMB=1.0, VX_AMB=0.01, VY_AMB=-0.02, VZ_AMB=0.03,
                   NOZZLE_PERTURB="random", TURB_VAMP=0.02, TURB_PAMP=0.1, N_TRACERS=ntr, TRACER_OFFSET=nvb,
                   TRACER_NOZZLE_VALUES=[1.0, 0.5, 0.2][:ntr], TRACER_AMB_VALUES=[0.0, 0.25, 0.1][:ntr],
                   TWO_TEMPERATURE=tt, THERMO_OFFSET=nvb + ntr, TE_NOZZLE=0.7, TI_NOZZLE=0.9,
                   CHEMISTRY_ENABLED=True, CHEM_OFFSET=coff, CHEM_X_HII_NOZZLE=0.9, CHEM_X_HEII_NOZZLE=0.3,
                   CHEM_X_HEIII_NOZZLE=0.05, PHYSICS=phys, B_INIT=binit, B0=0.5)
        pr = source_state(nshape, 760 + c, nv, chem_off=coff, bad=False)
        arrs[f"noz_in_{c}"] = pr.copy()
        rng = np.random.default_rng(1234)
        nozzle.apply_nozzle_left_x(pr, 1.0 / NX, 1.0 / NY, 1.0 / NZ, NY, NZ, 0.45, 0.55, rng, cfg)
        arrs[f"noz_out_{c}"] = pr[:, :ng]  # the fill writes the -x ghost planes only
    _save("sources", cool_cases=np.array([repr(c) for c in COOLING_CASES]),
          chem_cases=np.array([repr(c) for c in CHEM_CASES]),
          noz_cases=np.array(["/".join(map(str, c)) for c in noz_cases]), ng=np.array(ng), **arrs)


# ---------------------------------------------------------------------------
# Israel-Stewart dissipation: the SRHD RHS hooks (core/srhd_core.py:393-462), the
# step with fields 5..14 carried (PASSIVE_OFFSET = 15) and the post-step operator
# (core/dissipation.py:5-215)
# ---------------------------------------------------------------------------
DISS_RHS_CASES = [("muscl", "mc", "hllc"), ("muscl", "minmod", "hlle"), ("ppm", "mc", "hlle"), ("weno", "mc", "hllc")]
DISS_STEP_CASES = [("muscl", "vanleer", "hllc", 3), ("ppm", "mc", "hlle", 2)]
DISS_OP_CASES = [
    (True, dict(BULK_ZETA=0.05, SHEAR_ETA=0.02, HEAT_KAPPA=0.01, RELAX_TAU=0.1)),
    (True, dict(BULK_ZETA=0.05, RELAX_TAU_BULK=0.05, DISSIPATION_ADVECT=False)),
    (True, dict(SHEAR_ETA=5.0, HEAT_KAPPA=2.0, RELAX_TAU=0.01, DISSIPATION_CAP_FRAC=0.002, P_MAX=2.0)),
    (True, dict(BULK_ZETA=0.0, SHEAR_ETA=0.0, HEAT_KAPPA=0.0)),
    (True, dict(BULK_ZETA=0.1, HEAT_KAPPA=0.3, RELAX_TAU_SHEAR=0.0, RELAX_TAU=0.2, RELAX_TAU_HEAT=None)),
    (False, dict(BULK_ZETA=0.1, SHEAR_ETA=0.1, HEAT_KAPPA=0.1)),
]


def diss_state(shape, seed, npass=2):
    """srhd_state with the ten dissipative fields (Pi, pi^ij, q^i ~ +-4e-3) at 5..14 and the
    passives after them (PASSIVE_OFFSET = 15)."""
    base = srhd_state(shape, seed, npass=npass)
    diss = np.random.default_rng(seed + 1).uniform(-0.004, 0.004, (10,) + shape)
    return np.concatenate([base[:5], diss, base[5:]])


def gen_dissipation():
    # one shared input per group keeps the fixture small: 10^3 cells (4^3 interior with NG = 3)
    shape = (10, 10, 10)
    dx, dy, dz = 1.0 / 4, 1.0 / 5, 1.0 / 6
    arrs = {}
    pr = diss_state(shape, 900)
    arrs["rhs_in"] = pr
    for c, (recon, lim, riem) in enumerate(DISS_RHS_CASES):
        srhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, N_PASSIVE=2, PASSIVE_OFFSET=15)
        arrs[f"rhs_{c}"] = srhd_rhs(pr, recon, dx, dy, dz)
        print(f"[golden] dissipative rhs {recon}/{lim}/{riem}", flush=True)
    dt = 0.02
    pr = diss_state(shape, 920)
    arrs["step_in"] = pr
    for c, (recon, lim, riem, rk) in enumerate(DISS_STEP_CASES):
        srhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, N_PASSIVE=2, PASSIVE_OFFSET=15)
        arrs[f"step_{c}"] = (srhd_core.step_ssprk3 if rk == 3 else srhd_core.step_ssprk2)(pr, dx, dy, dz, dt)
        print(f"[golden] dissipative step {recon} rk{rk}", flush=True)
    op_in = diss_state(shape, 940, npass=0)
    arrs["op_in"] = op_in
    for c, (enabled, cfg) in enumerate(DISS_OP_CASES):
        full = dict(cfg, DISSIPATION_ENABLED=enabled, PHYSICS="hydro", NG=3)
        out = dissipation.apply_causal_dissipation(op_in.copy(), 0.01, dx, dy, dz, full)
        arrs[f"op_{c}"] = out[5:15]
    _save("dissipation", rhs_cases=np.array(["/".join(c) for c in DISS_RHS_CASES]),
          step_cases=np.array(["/".join(map(str, c)) for c in DISS_STEP_CASES]),
          op_cases=np.array([repr(c) for c in DISS_OP_CASES]), dxyz=np.array([dx, dy, dz]), dt=np.array(dt),
          **arrs)


# ---------------------------------------------------------------------------
# GRHD / GRMHD wrappers (core/grhd_core.py, core/grmhd_core.py, core/gr_metric.py)
# ---------------------------------------------------------------------------
GR_RHS_CASES = [
    # physics, GR_METRIC, ORTHONORMAL_FLUX, GR_MASS, GAMMA, recon, limiter, riemann
    ("grhd", "schwarzschild", True, 0.05, 1.4, "muscl", "mc", "hllc"),
    ("grhd", "kerr-schild", False, 0.05, 5.0 / 3.0, "ppm", "mc", "hlle"),
    ("grhd", "kerr-schild", True, 0.1, 5.0 / 3.0, "weno", "mc", "hllc"),
    ("grhd", "kerr", True, 0.05, 5.0 / 3.0, "muscl", "vanleer", "hlle"),
    ("grhd", "minkowski", True, 0.05, 5.0 / 3.0, "muscl", "mc", "hllc"),
    ("grmhd", "schwarzschild", True, 0.05, 4.0 / 3.0, "weno", "mc", "hlld_full"),
    ("grmhd", "kerr-schild", True, 0.05, 4.0 / 3.0, "muscl", "minmod", "hlld"),
]
GR_STEP_CASES = [  # index into GR_RHS_CASES, RK order
    (0, 3), (5, 2),
]


def gr_configure(phys, metric, ortho, mass, gamma, recon, lim, riem):
    gcfg = dict(GR_METRIC=metric, ORTHONORMAL_FLUX=ortho, GR_MASS=mass)
    if phys == "grhd":
        srhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, GAMMA=gamma)
        grhd_core.configure(gcfg)
        return grhd_core
    rmhd_configure(RECON=recon, LIMITER=lim, RIEMANN=riem, GAMMA=gamma)
    grmhd_core.configure(gcfg)
    return grmhd_core


def gen_gr():
    shape = (10, 11, 12)
    ng = 3
    dx, dy, dz = 1.0 / 4, 1.0 / 5, 1.0 / 6
    arrs = {}
    for c, case in enumerate(GR_RHS_CASES):
        core = gr_configure(*case)
        pr = (rmhd_state if case[0] == "grmhd" else srhd_state)(shape, 960 + c, jet=False)
        arrs[f"rhs_in_{c}"] = pr
        f = core.compute_rhs_grhd if case[0] == "grhd" else core.compute_rhs_grmhd
        arrs[f"rhs_{c}"] = f(pr, dx, dy, dz, 0, ng)
        print(f"[golden] gr rhs {case}", flush=True)
    dt = 0.02
    for c, (ci, rk) in enumerate(GR_STEP_CASES):
        core = gr_configure(*GR_RHS_CASES[ci])
        pr = arrs[f"rhs_in_{ci}"]
        arrs[f"step_{c}"] = (core.step_ssprk3 if rk == 3 else core.step_ssprk2)(pr, dx, dy, dz, dt, 0, ng)
        print(f"[golden] gr step {GR_RHS_CASES[ci]} rk{rk}", flush=True)
    _save("gr", rhs_cases=np.array([repr(c) for c in GR_RHS_CASES]), step_cases=np.array(GR_STEP_CASES),
          dxyz=np.array([dx, dy, dz]), dt=np.array(dt), ng=np.array(ng), **arrs)


# ---------------------------------------------------------------------------
# Driver loop (restates the loop order of solvers/srhd3d_mpi_muscl.py:276-326, one rank)
# ---------------------------------------------------------------------------
def run_driver(s, pr, nsteps, rng):
    phys = s["PHYSICS"]
    core = rmhd_core if phys in ("rmhd", "grmhd") else srhd_core
    core.configure(s)
    gr = {"grhd": grhd_core, "grmhd": grmhd_core}.get(phys)
    if gr is not None:
        gr.configure(s)
    ng = s["NG"]
    dx, dy, dz = s["Lx"] / s["NX"], s["Ly"] / s["NY"], s["Lz"] / s["NZ"]
    y0, z0 = s["JET_CENTER"][1], s["JET_CENTER"][2]
    boundary.apply_periodic_yz(pr, ng)
    boundary.apply_outflow_right_x(pr, ng)
    nozzle.apply_nozzle_left_x(pr, dx, dy, dz, s["NY"], s["NZ"], y0, z0, rng, s)
    dts = []
    for step in range(nsteps):
        if s["PHYSICS"] in ("hydro", "grhd"):
            amax = srhd_core.max_char_speed(pr, *pr.shape[1:])
        else:
            amax = 1.0
        dt = s["CFL"] * min(dx, dy, dz) / max(amax, 1e-12)
        boundary.apply_periodic_yz(pr, ng)
        boundary.apply_outflow_right_x(pr, ng)
        nozzle.apply_nozzle_left_x(pr, dx, dy, dz, s["NY"], s["NZ"], y0, z0, rng, s)
        if gr is not None:
            pr = (gr.step_ssprk3 if s["RK_ORDER"] == 3 else gr.step_ssprk2)(pr, dx, dy, dz, dt, 0, ng)
        else:
            pr = (core.step_ssprk3 if s["RK_ORDER"] == 3 else core.step_ssprk2)(pr, dx, dy, dz, dt)
        boundary.apply_periodic_yz(pr, ng)
        boundary.apply_outflow_right_x(pr, ng)
        nozzle.apply_nozzle_left_x(pr, dx, dy, dz, s["NY"], s["NZ"], y0, z0, rng, s)
        pr = dissipation.apply_causal_dissipation(pr, dt, dx, dy, dz, s)
        pr = source_terms.apply_cooling_heating(pr, dt, s)
        pr = source_terms.apply_two_temperature(pr, dt, s)
        pr = chemistry.apply_ion_chemistry(pr, dt, s)
        dts.append(dt)
        print(f"[golden] driver step {step+1}/{nsteps} dt={dt:.4e} amax={amax:.4f}", flush=True)
    return pr, np.array(dts)


def driver_settings(**kw):
    s = dict(NX=8, NY=8, NZ=8, Lx=1.0, Ly=1.0, Lz=1.0, NG=2, CFL=0.4, GAMMA=5.0 / 3.0,
             JET_RADIUS=0.125, JET_CENTER=[0.0, 0.5, 0.5], GAMMA_JET=7.0888121, ETA_RHO=0.01, P_EQ=0.01,
             SHEAR_THICK=0.0, NOZZLE_PROFILE="top_hat", NOZZLE_PERTURB="none", TURB_VAMP=0.02, TURB_PAMP=0.0,
             RHO_AMB=1.0, P_AMB=0.01, VX_AMB=0.0, VY_AMB=0.0, VZ_AMB=0.0, P_MAX=2.0, V_MAX=0.9999,
             RECON="muscl", LIMITER="minmod", RIEMANN="hlle", RK_ORDER=2, PHYSICS="hydro",
             GLM_CH=1.0, GLM_CP=0.1, B_INIT="none", B0=0.0, N_TRACERS=0, TRACER_OFFSET=5,
             N_PASSIVE=0, PASSIVE_OFFSET=5, TWO_TEMPERATURE=False, TEI_TAU=0.1, THERMO_OFFSET=5,
             RESISTIVE_ENABLED=False, RESISTIVITY=0.0)
    s.update(kw)
    return s


def init_block(s, noise_seed):
    nv = (9 if s["PHYSICS"] in ("rmhd", "grmhd") else (15 if s.get("DISSIPATION_ENABLED", False) else 5)) + \
        s["N_PASSIVE"]
    ng = s["NG"]
    pr = np.zeros((nv, s["NX"] + 2 * ng, s["NY"] + 2 * ng, s["NZ"] + 2 * ng))
    pr[0] = s["RHO_AMB"]; pr[4] = s["P_AMB"]
    if s["TWO_TEMPERATURE"]:
        pr[s["THERMO_OFFSET"]] = s["TE_AMB"]; pr[s["THERMO_OFFSET"] + 1] = s["TI_AMB"]
    if s.get("CHEMISTRY_ENABLED", False):
        for m, key in enumerate(("CHEM_X_HII_AMB", "CHEM_X_HEII_AMB", "CHEM_X_HEIII_AMB")):
            pr[s["CHEM_OFFSET"] + m] = s[key]
    if noise_seed is not None:
        noise = np.random.default_rng(noise_seed).uniform(-0.01, 0.01, (s["NX"], s["NY"], s["NZ"]))
        pr[0, ng:-ng, ng:-ng, ng:-ng] = s["RHO_AMB"] * (1.0 + noise)
    return pr


def gen_driver_small():
    # RMHD WENO5 + hlld_full + SSPRK3, NG = 3, random nozzle perturbation, toroidal nozzle field,
    # two-temperature passives, resistivity: exercises the whole loop order on a tiny grid.
    # a mildly relativistic jet (v = 0.3): the reference RMHD step diverges for W ~ 7 jets
    s = driver_settings(NX=8, NY=8, NZ=8, NG=3, PHYSICS="rmhd", GAMMA=4.0 / 3.0, RECON="weno",
                        RIEMANN="hlld_full", RK_ORDER=3, B_INIT="toroidal", B0=0.05, NOZZLE_PERTURB="random",
                        TURB_PAMP=0.05, JET_RADIUS=0.3, GAMMA_JET=1.0482848, ETA_RHO=0.5, SHEAR_THICK=0.05,
                        TWO_TEMPERATURE=True, TEI_TAU=0.1, TE_AMB=0.01,
                        TI_AMB=0.01, TE_NOZZLE=0.5, TI_NOZZLE=0.2, N_PASSIVE=2, PASSIVE_OFFSET=9,
                        THERMO_OFFSET=9, TRACER_OFFSET=9, RESISTIVE_ENABLED=True, RESISTIVITY=1e-3)
    pr0 = init_block(s, 12345)
    pr0[5] = 0.1  # axial field everywhere (as the RMHD benchmark's B_AMB)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 3, np.random.default_rng(1234))
    print(f"[golden] rmhd driver {time.time()-t0:.1f}s: rho [{out[0].min():.3e}, {out[0].max():.3e}] "
          f"p max {out[4].max():.3e} |B| max {np.abs(out[5:8]).max():.3e}", flush=True)
    _save("driver_rmhd_8", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))

    s = driver_settings(NX=10, NY=9, NZ=8, NG=2, RECON="muscl", LIMITER="mc", RIEMANN="hllc", RK_ORDER=3,
                        NOZZLE_PERTURB="random", TURB_PAMP=0.05, SHEAR_THICK=0.02, JET_RADIUS=0.3,
                        N_TRACERS=1, N_PASSIVE=1, TRACER_NOZZLE_VALUES=[1.0], TRACER_AMB_VALUES=[0.0])
    pr0 = init_block(s, 777)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 3, np.random.default_rng(1234))
    print(f"[golden] srhd small driver {time.time()-t0:.1f}s", flush=True)
    _save("driver_srhd_10", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))


def gen_driver_sources():
    # SRHD MUSCL-MC + HLLC + SSPRK3 jet with every post-step source on: cooling/heating,
    # two-temperature and H/He chemistry (rates and energy coupling scaled up so that all of them
    # move the state within 3 steps), a tracer, and the nozzle's ion fractions.
    s = driver_settings(NX=10, NY=9, NZ=8, NG=2, RECON="muscl", LIMITER="mc", RIEMANN="hllc", RK_ORDER=3,
                        NOZZLE_PERTURB="random", TURB_PAMP=0.05, SHEAR_THICK=0.02, JET_RADIUS=0.3,
                        N_TRACERS=1, TRACER_NOZZLE_VALUES=[1.0], TRACER_AMB_VALUES=[0.0], TRACER_OFFSET=5,
                        TWO_TEMPERATURE=True, TEI_TAU=0.1, THERMO_OFFSET=6, TE_AMB=0.01, TI_AMB=0.01,
                        TE_NOZZLE=0.5, TI_NOZZLE=0.2,
                        COOLING_ENABLED=True, COOLING_LAMBDA=0.5, HEATING_RATE=0.005,
                        CHEMISTRY_ENABLED=True, N_CHEM=3, CHEM_OFFSET=8, CHEM_TUNIT_K=1.0e6,
                        CHEM_RATE_SCALE=1.0e8, CHEM_ENERGY_SCALE=1.0e-3, CHEM_X_HII_AMB=0.1,
                        CHEM_X_HEII_AMB=0.05, CHEM_X_HEIII_AMB=0.0, CHEM_X_HII_NOZZLE=0.9,
                        CHEM_X_HEII_NOZZLE=0.3, CHEM_X_HEIII_NOZZLE=0.05, N_PASSIVE=6, PASSIVE_OFFSET=5)
    pr0 = init_block(s, 4242)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 3, np.random.default_rng(1234))
    print(f"[golden] srhd sources driver {time.time()-t0:.1f}s: xHII [{out[8].min():.3e}, {out[8].max():.3e}]",
          flush=True)
    _save("driver_sources_10", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))


def gen_driver_dissipation():
    # SRHD MUSCL-MC + HLLC + SSPRK3 jet with Israel-Stewart dissipation (fields 5..14, a tracer at 15),
    # bulk, shear and heat terms on, upwind advection of the fields, perturbed nozzle.
    s = driver_settings(NX=10, NY=9, NZ=8, NG=2, RECON="muscl", LIMITER="mc", RIEMANN="hllc", RK_ORDER=3,
                        NOZZLE_PERTURB="random", TURB_PAMP=0.05, SHEAR_THICK=0.02, JET_RADIUS=0.3,
                        DISSIPATION_ENABLED=True, BULK_ZETA=0.01, SHEAR_ETA=0.005, HEAT_KAPPA=0.002,
                        RELAX_TAU=0.1, DISSIPATION_CAP_FRAC=0.5, DISSIPATION_ADVECT=True,
                        N_TRACERS=1, TRACER_NOZZLE_VALUES=[1.0], TRACER_AMB_VALUES=[0.0], TRACER_OFFSET=15,
                        N_PASSIVE=1, PASSIVE_OFFSET=15)
    pr0 = init_block(s, 5151)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 3, np.random.default_rng(1234))
    print(f"[golden] srhd dissipation driver {time.time()-t0:.1f}s: |Pi| max {np.abs(out[5]).max():.3e} "
          f"|pi_xy| max {np.abs(out[9]).max():.3e} |q_x| max {np.abs(out[12]).max():.3e}", flush=True)
    _save("driver_diss_10", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))


def gen_driver_gr():
    # GRHD jet, MUSCL-MC + HLLC + SSPRK3, Schwarzschild metric (isotropic lapse sources, Kerr-Schild
    # lapse in the flux velocities) with the hole at the domain corner, Gamma = 1.4 in the sources
    s = driver_settings(NX=10, NY=9, NZ=8, NG=2, PHYSICS="grhd", GAMMA=1.4, RECON="muscl", LIMITER="mc",
                        RIEMANN="hllc", RK_ORDER=3, NOZZLE_PERTURB="random", TURB_PAMP=0.05, SHEAR_THICK=0.02,
                        JET_RADIUS=0.3, GAMMA_JET=2.0, GR_METRIC="schwarzschild", GR_MASS=0.05,
                        ORTHONORMAL_FLUX=True)
    pr0 = init_block(s, 6161)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 3, np.random.default_rng(1234))
    print(f"[golden] grhd driver {time.time()-t0:.1f}s: rho [{out[0].min():.3e}, {out[0].max():.3e}]", flush=True)
    _save("driver_grhd_10", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))


def gen_driver_32():
    # The 32^3 SRHD jet: MUSCL-minmod + HLLE + SSPRK2, CFL 0.4, 20 steps, +-1% density noise
    # (default_rng(12345)), nozzle radius 4 cells, W_jet = 7.0888; reference periodic y/z boundaries.
    s = driver_settings(NX=32, NY=32, NZ=32, NG=2, RECON="muscl", LIMITER="minmod", RIEMANN="hlle", RK_ORDER=2)
    pr0 = init_block(s, 12345)
    t0 = time.time()
    out, dts = run_driver(s, pr0.copy(), 20, np.random.default_rng(1234))
    print(f"[golden] srhd 32^3 driver {time.time()-t0:.1f}s", flush=True)
    _save("driver_srhd_32", pr0=pr0, out=out, dts=dts, settings=np.array(repr(s)))


GENS = dict(pointwise=gen_pointwise, rhs=gen_rhs, rhs_passive=gen_rhs_passive, steps=gen_steps, bc=gen_bc, sources=gen_sources,
            driver_small=gen_driver_small, driver_sources=gen_driver_sources,
            dissipation=gen_dissipation, driver_diss=gen_driver_dissipation, gr=gen_gr, driver_gr=gen_driver_gr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    ap.add_argument("--driver", action="store_true")
    a = ap.parse_args()
    names = a.only if a.only else list(GENS)
    for n in names:
        t0 = time.time()
        GENS[n]()
        print(f"[golden] {n} done in {time.time()-t0:.1f}s", flush=True)
    if a.driver:
        gen_driver_32()


if __name__ == "__main__":
    main()
