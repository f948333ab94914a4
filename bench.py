#!/usr/bin/env python3
"""Benchmark of the SRHD/RMHD finite-volume step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmhd_jet_256]

Workload (default): the RMHD 256^3 jet, WENO5 + HLLD + SSPRK3, Gamma = 4/3
(astrofd_amd/config/rmhd_jet_256.json).  A step is one pass of the hot path over
the whole grid, as the reference driver runs it: ghost/nozzle fills, one SSPRK3
step, ghost/nozzle fills.  Inputs are resident in HBM before the timed region.

Timed runs from the same seeded initial state:
  plain  FP64 everything, plain con2prim (the baseline of the metric)
  mixed  FP32 reconstruction + Riemann fluxes, FP64 Lorentz factor / con2prim /
         SSPRK accumulation, learned con2prim guess when weights are available
  mixed with the plain con2prim (what the learned guess changes in the step)
  mixed with the previous-state Newton start (RMHD)
  mixed with one thread per face (--face-path-ab: what the shared-edge face kernels change)
`value` is the mixed run's zone-updates/s.  A lockstep pass of plain and mixed
reports the relative L1 and max error of rho, p and W after checkpoint steps.

Prints ONE JSON line.  Also: the dominant kernel's roofline from HIP events on
the library's stream over the timed steps, the solver counters of one extra
step, the fused post-step source kernel (cooling, two-temperature, chemistry)
on the same grid, and the CPU oracle's rate on a bounded sample on this host's
cores.  Beside the headline:
  fp64_ulp_spread  the plain FP64 step's own response to a 1-ulp perturbation of its input
                   (the reference's hll_d_flux_dir is ill-conditioned over a static ambient);
                   every errors entry carries the same spread after its own step
  hlle_50          50 consecutive steps of the same jet with RIEMANN = "hlle" (the reference's
                   own WENO RMHD choice, stable on this jet): mixed + learned against plain FP64
  network          the learned guess's MLP alone over 2^24 samples: achieved BF16 MFMA rate
  srhd_jet_256     the SRHD 256^3 jet (MUSCL-van Leer + HLLC + SSPRK3): timed plain and mixed
                   steps, errors after steps 1, 10 and 100, and the cells whose con2prim took the
                   other branch (converged vs energy fallback) in the mixed run
  rmhd_jet_512     the 512^3 RMHD jet with resistivity and Te/Ti: timed plain and mixed + learned
                   steps and the errors after one step
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
MLP_WEIGHTS = os.path.join(ROOT, "astrofd_amd", "models", "c2p_mlp.npz")
DEFAULT_WORKLOAD = "rmhd_jet_256"  # the workload the headline metric is quoted on


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_workload(name):
    from astrofd_amd.settings import finalize, load_config_file
    from astrofd_amd import ic
    s = finalize(load_config_file(os.path.join(ROOT, "astrofd_amd", "config", name + ".json")))
    ic.apply_derived_ic_keys(s)
    return s


def lorentz(pr):
    v2 = pr[1] ** 2 + pr[2] ** 2 + pr[3] ** 2
    return 1.0 / np.sqrt(1.0 - np.minimum(v2, 1.0 - 1e-14))


def errors(a, b, ng):
    """Relative L1 and max |diff| of rho, p, W over interior cells (b is the reference), over the
    cells where both states are finite."""
    sl = (slice(ng, -ng),) * 3
    out = {}
    for name, fa, fb in (("rho", a[0][sl], b[0][sl]), ("p", a[4][sl], b[4][sl]),
                         ("W", lorentz(a)[sl], lorentz(b)[sl])):
        ok = np.isfinite(fa) & np.isfinite(fb)
        d = np.abs(fa[ok] - fb[ok])
        ref = np.abs(fb[ok]).sum()
        out[name] = {"L1_rel": float(d.sum() / ref) if ref > 0 else 0.0,
                     "max_abs": float(d.max()) if d.size else 0.0, "cells": int(ok.sum())}
    return out


def sync(d):
    d.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def make_stepper(d, s, rng):
    dx = min(s["Lx"] / s["NX"], s["Ly"] / s["NY"], s["Lz"] / s["NZ"])

    def one_step():
        # the reference driver's step: CFL (SRHD only; RMHD uses amax = 1), fills, SSPRK, fills, sources
        amax = d.max_char_speed() if s["PHYSICS"] in ("hydro", "grhd") else 1.0
        dt = s["CFL"] * dx / max(amax, 1e-12)
        d.apply_bcs(rng)
        d.step(dt)
        d.apply_bcs(rng)
        d.post_step_sources(dt)  # none in the RMHD jet workload: returns without a launch
    return one_step


def open_domain(s, pr0, mlp=None, face_path=0):
    from astrofd_amd.domain import Domain
    d = Domain.from_settings(s, face_path=face_path)
    d.set_prim(pr0)
    d.set_nozzle(s)
    if mlp is not None:
        d.set_mlp_weights(mlp)
    rng = np.random.default_rng(1234)
    d.apply_bcs(rng)
    return d, rng


def run_timed(s, pr0, steps, warmup, mlp=None, repeat=True, face_path=0):
    """Time `steps` steps after `warmup` untimed ones, with per-kernel HIP events on the library's
    stream, then one untimed step with the solver counters on.

    repeat=True: every step starts from the seeded initial state (re-uploaded and filled
    outside the timed interval), so each timed step is one pass of the hot path over the same
    synthetic input.  SRHD: the reset also leaves the input's CFL maximum over the speed box, as
    a run's previous step does, so the timed step does a run's step's CFL work (the ghost shell
    before it, the next step's box maximum in its last stage).  repeat=False: consecutive steps
    of one run."""
    d, rng = open_domain(s, pr0, mlp, face_path)
    try:
        one_step = make_stepper(d, s, rng)

        def reset():
            if repeat:
                d.set_prim(pr0)
                d.apply_bcs(rng)
                if s["PHYSICS"] in ("hydro", "grhd"):
                    # the speed box's maximum of the input, as the previous step's last stage leaves
                    # it in a run; the timed step adds the ghost shell and takes the next step's
                    # maximum in its own last stage
                    d.max_char_speed()

        for _ in range(warmup):
            reset()
            one_step()
        d.set_instrumentation(stats=False, timing=True)
        d.reset_stats()
        elapsed = 0.0
        if repeat:
            for _ in range(steps):
                d.set_instrumentation(stats=False, timing=False)
                reset()
                d.set_instrumentation(stats=False, timing=True)
                sync(d)
                t0 = time.perf_counter()
                one_step()
                sync(d)
                elapsed += time.perf_counter() - t0
        else:
            sync(d)
            t0 = time.perf_counter()
            for _ in range(steps):
                one_step()
            sync(d)
            elapsed = time.perf_counter() - t0
        kern = d.stats()
        d.set_instrumentation(stats=True, timing=False)
        d.reset_stats()
        reset()
        one_step()
        st = d.stats()
        d.set_instrumentation(stats=False, timing=False)
        return dict(elapsed=elapsed, kernels=kern, stats=st)
    finally:
        d.close()


def ulp_perturbed(pr0):
    """The state with every value moved by 1 ulp (seeded signs)."""
    return pr0 * (1.0 + 2.220446049250313e-16 * np.random.default_rng(99).choice([-1.0, 1.0], pr0.shape))


def accuracy(sp, sm, pr0, checkpoints, mlp=None, flips=False):
    """Plain and mixed runs in lockstep from the same state; errors of mixed against plain after
    each checkpoint step (cells where both are finite), and the non-finite counts.  A third run,
    plain FP64 from the state perturbed by 1 ulp, gives the scheme's own sensitivity at the same
    steps (fp64_ulp_spread: its errors against the plain run): an implementation that rounds
    differently from the plain run cannot be expected to stay closer to it than that.  flips:
    also count the cells whose last-stage con2prim took the other branch of the hydro fixed point
    (converged vs energy fallback) in the two runs, at each checkpoint step."""
    dp, rp = open_domain(sp, pr0)
    dm, rm = open_domain(sm, pr0, mlp)
    dq, rq = open_domain(sp, ulp_perturbed(pr0))
    out = []
    try:
        step_p, step_m, step_q = make_stepper(dp, sp, rp), make_stepper(dm, sm, rm), make_stepper(dq, sp, rq)
        for n in range(1, max(checkpoints) + 1):
            cp = n in checkpoints
            if cp and flips:
                dp.set_instrumentation(stats=True, timing=False)
                dm.set_instrumentation(stats=True, timing=False)
            step_p()
            step_m()
            step_q()
            if cp:
                a, b = dm.get_prim(), dp.get_prim()
                e = errors(a, b, sp["NG"])
                e["step"] = n
                e["nonfinite"] = {"plain": int(np.count_nonzero(~np.isfinite(b))),
                                  "mixed": int(np.count_nonzero(~np.isfinite(a)))}
                # the largest density of each run: where the scheme's own solution has left the
                # range an FP32 flux can carry (|F| < 3.4e38), the mixed path's fluxes overflow
                ng = sp["NG"]
                e["rho_max"] = {"plain": float(np.nanmax(b[0][(slice(ng, -ng),) * 3])),
                                "mixed": float(np.nanmax(np.where(np.isfinite(a[0]), a[0], np.nan)[(slice(ng, -ng),) * 3]))}
                del a
                q = dq.get_prim()
                e["fp64_ulp_spread"] = {k: {"L1_rel": v["L1_rel"], "max_abs": v["max_abs"]}
                                        for k, v in errors(q, b, sp["NG"]).items()}
                del q, b
                if flips:
                    ng = sp["NG"]
                    sl = (slice(ng, -ng),) * 3
                    fp, fm = dp.c2p_flags()[sl] & 1, dm.c2p_flags()[sl] & 1
                    e["c2p_branch_flips"] = int(np.count_nonzero(fp != fm))
                    e["c2p_fallback_cells"] = {"plain": int(fp.sum()), "mixed": int(fm.sum())}
                    dp.set_instrumentation(stats=False, timing=False)
                    dm.set_instrumentation(stats=False, timing=False)
                out.append(e)
                log(f"[bench]   accuracy after step {n}: L1(rho) {e['rho']['L1_rel']:.3e}, nonfinite {e['nonfinite']}"
                    + (f", con2prim branch flips {e['c2p_branch_flips']}" if flips else ""))
    finally:
        dp.close()
        dm.close()
        dq.close()
    return out


def measured_traffic(workload):
    """HBM bytes per launch by kernel, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/<workload>_traffic.txt, scaled by a plain-copy calibration), and its source, or
    ({}, None)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.txt")
    if not os.path.exists(path):
        return {}, None
    with open(path) as f:
        doc = json.load(f)
    src = f"profiles/{workload}_traffic.txt: " + doc.get("source", "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE")
    return doc.get("traffic_bytes_per_launch", {}), src


def face_variant(s, mixed, face_path=0):
    """The mixed RMHD WENO5 face-kernel variant a run takes (afd_params.face_path; 0 selects
    AFD_FACE_PATH_DEFAULT): 1 one thread per face, 2 shared edges with stored face fluxes,
    3 flux divergence, 4 flux divergence with the y faces in the stage update.  0 for every other
    workload (one thread per face)."""
    from astrofd_amd import _lib as L
    from astrofd_amd.domain import recon_id
    cells = (s["NX"] + 2 * s["NG"]) * (s["NY"] + 2 * s["NG"]) * (s["NZ"] + 2 * s["NG"])
    if not (s["PHYSICS"] == "rmhd" and mixed and recon_id(s["RECON"]) == 2):
        return 0
    v = face_path or L.AFD_FACE_PATH_DEFAULT
    return v if v == 1 or cells < (1 << 29) else 1


def flux_divergence_path(s, mixed, face_path=0):
    """The face kernels accumulate the flux divergence (astrofd_amd/csrc/afd_kernels.hpp,
    k_face_rmhd_weno_lanes_div / _march) instead of storing face fluxes."""
    return face_variant(s, mixed, face_path) >= 3


def kernel_name(s, which, mixed, learned, face_path=0):
    """The rocprofv3 name of a kernel of this workload (face_x/y/z or stage_update_c2p)."""
    T = "float" if mixed else "double"
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    v = face_variant(s, mixed, face_path)
    from astrofd_amd.domain import recon_id, riemann_id
    rid = riemann_id(s['PHYSICS'], s['RIEMANN'])
    if which == "stage_update_c2p":
        if v == 4:
            return f"k_face_march_update<{rid}, {'true' if learned else 'false'}>"
        return f"k_update<{nb}, {T}, {'true' if learned else 'false'}>"
    d = {"face_x": 0, "face_y": 1, "face_z": 2}[which]
    if v >= 3:
        # FP32 WENO5 faces accumulating the flux divergence (pencils for x/y, lanes for z)
        return f"k_face_rmhd_weno_lanes_div<{rid}>" if d == 2 else f"k_face_rmhd_weno_march<{rid}, {d}>"
    if v == 2:
        # FP32 WENO5 faces sharing edge values (LDS rows for x/y, lanes for z), fluxes stored
        return f"k_face_rmhd_weno_lanes<{rid}>" if d == 2 else f"k_face_rmhd_weno_rows<{rid}, {d}>"
    kind = "k_face_rmhd" if nb == 9 else "k_face_srhd"
    return f"{kind}<{recon_id(s['RECON'])}, {rid}, {d}, {T}>"


def nvars(s):
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    return nb + int(s.get("N_PASSIVE", 0))


def algorithmic_update_bytes(s):
    """Compulsory HBM bytes per cell of one stage update as the metric counts them: an ideally
    fused SSPRK stage reads P_s-1, U0 and U_s-1 and writes U_s and P_s, 13 state arrays a step for
    SSPRK3 and 8 for SSPRK2, each nv x 8 B, spread over the stages."""
    nst = 3 if s["RK_ORDER"] == 3 else 2
    return (13 if nst == 3 else 8) * nvars(s) * 8 / nst


def impl_update_bytes(s, mixed, face_path=0):
    """HBM bytes per cell of one stage update as this implementation moves them, averaged over the
    stages: read P0 (U0 = p2c(P0) every stage), read U_s-1 (stages > 0), write U_s (not the last
    stage), write P_s, read psi and the passives of P_s-1, and the face terms: one FP32 flux
    divergence on the flux-divergence path, else one read of each face-flux array."""
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    nv = nvars(s)
    nst = 3 if s["RK_ORDER"] == 3 else 2
    fb = 4 if mixed else 8
    carried = nb == 9 and s["RECON"] != "muscl"
    face = nb * 4 if flux_divergence_path(s, mixed, face_path) else (4 if carried else 3) * nb * fb
    per = []
    for st in range(nst):
        b = nv * 8 + nv * 8 + face + (8 if nb == 9 else 0) + (nv - nb) * 8
        b += nb * 8 if st > 0 else 0
        b += nb * 8 if st < nst - 1 else 0
        per.append(b)
    return sum(per) / nst


def face_bytes(s, mixed, dkey, face_path=0):
    """Compulsory bytes per face (flux path) or per cell (flux-divergence path) of a face kernel:
    read the nb primitives of one cell (the stencil's other cells are re-reads) and write its output:
    the face flux (twice for the RMHD PPM/WENO x faces, which also read the two carried speeds the
    z kernel wrote), or read and write the FP32 divergence (the z kernel writes it and the speeds)."""
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    fb = 4 if mixed else 8
    carried = nb == 9 and s["RECON"] != "muscl"
    if flux_divergence_path(s, mixed, face_path):
        return [nb * 8 + 2 * 4 + 2 * nb * 4, nb * 8 + 2 * nb * 4, nb * 8 + nb * 4 + 2 * 4][dkey]
    b = nb * 8 + nb * fb
    if carried and dkey == 0:
        b += nb * fb + 2 * fb
    if carried and dkey == 2:
        b += 2 * fb
    return b


def kernel_roofline(s, kern, mixed, learned=False, workload=None, face_path=0):
    """Dominant kernel and its achieved algorithmic HBM bandwidth (bytes per launch / avg duration).

    Stage update (one thread per cell, ghosts included): the metric's compulsory state traffic per
    stage (algorithmic_update_bytes) times the cells of a launch; impl_bytes_per_launch is what
    this implementation moves (impl_update_bytes).  Face kernels: face_bytes per face (or per cell
    on the flux-divergence path).  traffic: HBM bytes per launch measured by rocprofv3 in the
    committed profile named by traffic_source.
    """
    NX, NY, NZ = s["NX"], s["NY"], s["NZ"]
    full = (NX + 2 * s["NG"]) * (NY + 2 * s["NG"]) * (NZ + 2 * s["NG"])
    units = [(NX + 1) * NY * NZ, NX * (NY + 1) * NZ, NX * NY * (NZ + 1)]
    if flux_divergence_path(s, mixed, face_path):
        units = [NX * NY * NZ] * 3
    cands = []
    for dkey, name in ((0, "face_x"), (1, "face_y"), (2, "face_z")):
        n = kern["launches_face"][dkey]
        if n:
            nbytes = face_bytes(s, mixed, dkey, face_path) * units[dkey]
            cands.append((kern["ms_face"][dkey] / n, name, nbytes, nbytes))
    if kern["launches_update"]:
        cands.append((kern["ms_update"] / kern["launches_update"], "stage_update_c2p",
                      algorithmic_update_bytes(s) * full, impl_update_bytes(s, mixed, face_path) * full))
    ms, name, nbytes, impl = max(cands)
    achieved = nbytes / (ms * 1e-3) / 1e9
    kname = kernel_name(s, name, mixed, learned, face_path)
    table, src = measured_traffic(workload) if workload else ({}, None)
    traffic = table.get(kname)
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": round(traffic) if traffic else None,
            "traffic_source": src if traffic else None, "avg_ms": round(ms, 4), "bytes_per_launch": int(nbytes),
            "impl_bytes_per_launch": int(impl), "impl_frac": round(impl / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "rocprof_name": kname}


def cpu_baseline(s, n=128, steps=2, n1=48):
    """The CPU oracle (C restatement of the reference, FP64, OpenMP) on a bounded sample of the same
    workload: `steps` steps (fills + SSPRK step + fills) at n^3 on all host cores, and one step at
    n1^3 on one core (the reference itself is serial per rank)."""
    import oracle as O
    from astrofd_amd import ic
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass

    def run(m, nsteps, threads):
        t = dict(s)
        t.update(NX=m, NY=m, NZ=m)
        pr = ic.initial_state(t)
        rm = t["PHYSICS"] == "rmhd"
        P = O.make_params("rmhd" if rm else "hydro", pr.shape[1:], nv=pr.shape[0], ng=t["NG"],
                          gamma=t["GAMMA"] if rm else None, p_max=t["P_MAX"], v_max=t["V_MAX"],
                          glm_ch=t["GLM_CH"], glm_cp=t["GLM_CP"], limiter=t["LIMITER"], recon=t["RECON"],
                          riemann=t["RIEMANN"], rk_order=t["RK_ORDER"], n_passive=t["N_PASSIVE"])
        O.set_num_threads(threads)
        dx = 1.0 / m
        bamb = tuple(t.get("B_AMB") or (0.0, 0.0, 0.0))

        def bcs(a):
            O.apply_outflow_yz(P, a, t["BC_Y"] == "outflow", t["BC_Z"] == "outflow")
            O.apply_outflow_right_x(P, a)
            O.apply_nozzle_left_x(P, a, dx, dx, dx, m, m, t["JET_CENTER"][1], t["JET_CENTER"][2], t, None, bamb)

        bcs(pr)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            bcs(pr)
            amax = O.max_char_speed(P, pr) if not rm else 1.0
            pr = O.step(P, pr, dx, dx, dx, t["CFL"] * dx / amax)
            bcs(pr)
        return m ** 3 * nsteps / (time.perf_counter() - t0), time.perf_counter() - t0

    v_all, el_all = run(n, steps, cores)
    v_one, el_one = run(n1, 1, 1)
    return {"value": round(v_all, 1), "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "value_1core": round(v_one, 1),
            "sample": f"{steps} steps of the same workload at {n}^3 on {cores} threads ({el_all:.2f} s wall), "
                      f"and 1 step at {n1}^3 on 1 thread ({el_one:.2f} s): C oracle, FP64, OpenMP"}


SOURCE_BYTES_PER_CELL = 7 * 8 + 6 * 8  # read rho, p, Te, Ti, 3 ion fractions; write all but rho


def sources_roofline(s, pr0, launches=20):
    """The fused post-step source kernel (cooling/heating + two-temperature + H/He chemistry,
    solvers/srhd3d_mpi_muscl.py:324-326) on the workload's grid and state, with Te, Ti and the ion
    fractions appended as passives; HIP-event time per launch and its algorithmic HBM bandwidth
    (SOURCE_BYTES_PER_CELL per interior cell)."""
    from astrofd_amd.domain import Domain
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    t = dict(s, N_TRACERS=0, N_THERMO=2, N_CHEM=3, N_PASSIVE=5, PASSIVE_OFFSET=nb, THERMO_OFFSET=nb,
             CHEM_OFFSET=nb + 2, TWO_TEMPERATURE=True, TEI_TAU=0.5, COOLING_ENABLED=True, COOLING_LAMBDA=0.1,
             HEATING_RATE=1e-4, CHEMISTRY_ENABLED=True, CHEM_TUNIT_K=1e6, CHEM_RATE_SCALE=1e8,
             CHEM_ENERGY_SCALE=1e-6)
    pr = np.zeros((nb + 5,) + pr0.shape[1:])
    pr[:nb] = pr0[:nb]
    pr[nb] = pr[nb + 1] = pr0[4] / pr0[0]
    pr[nb + 2], pr[nb + 3] = 0.1, 0.05
    d = Domain.from_settings(t)
    try:
        d.set_prim(pr)
        del pr
        for _ in range(2):
            d.post_step_sources(1e-3)
        d.set_instrumentation(stats=False, timing=True)
        d.reset_stats()
        for _ in range(launches):
            d.post_step_sources(1e-3)
        sync(d)
        ms = d.stats()["ms_other"] / launches
        d.set_instrumentation(stats=False, timing=False)
        if d.count_nonfinite():
            raise RuntimeError("post-step sources produced non-finite values")
    finally:
        d.close()
    nbytes = SOURCE_BYTES_PER_CELL * s["NX"] * s["NY"] * s["NZ"]
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "post_step_sources", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "avg_ms": round(ms, 4),
            "bytes_per_launch": int(nbytes), "rocprof_name": "k_sources(SourceArgs)"}


MLP_FLOP_PER_SAMPLE = 2 * (5 * 64 + 64 * 64 + 64 * 2)  # 9,088: three dense layers, multiply + add
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def network_rate(mlp, n=1 << 24, launches=3):
    """The learned guess's MLP (5 -> 64 -> 64 -> 2, SiLU, BF16 weights, FP32 accumulation) alone over
    n samples: kernel time from HIP events on the library's stream, achieved FLOP rate against the
    dense BF16 MFMA peak.  The inputs are uniform in [-1, 1] (the rate does not depend on them)."""
    from astrofd_amd.domain import Domain
    s = dict(load_workload(DEFAULT_WORKLOAD), NX=8, NY=8, NZ=8, PRECISION="mixed", C2P_GUESS="learned")
    X = np.random.default_rng(5).uniform(-1.0, 1.0, (n, 5)).astype(np.float32)
    d = Domain.from_settings(s)
    try:
        d.set_mlp_weights(mlp)
        d.mlp_forward(X[:1024])
        d.set_instrumentation(stats=False, timing=True)
        d.reset_stats()
        for _ in range(launches):
            d.mlp_forward(X)
        ms = d.stats()["ms_other"] / launches
        d.set_instrumentation(stats=False, timing=False)
    finally:
        d.close()
    tflops = MLP_FLOP_PER_SAMPLE * n / (ms * 1e-3) / 1e12
    return {"kernel": "k_mlp_forward", "samples": n, "avg_ms": round(ms, 4), "flop_per_sample": MLP_FLOP_PER_SAMPLE,
            "achieved": round(tflops, 2), "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / BF16_MFMA_PEAK_TFLOPS, 5)}


def timed_entry(res, zu, steps):
    return {"value": round(zu / res["elapsed"], 1), "ms_per_step": round(res["elapsed"] * 1e3 / steps, 3),
            "kernel_ms_per_step": kernel_ms(res["kernels"], steps), "c2p": c2p_report(res["stats"])}


def hlle_leg(s, pr0, mlp, nsteps=50, checkpoints=(1, 10, 25, 50)):
    """The same jet with RIEMANN = "hlle" (stable on it), 50 consecutive steps: plain FP64 and mixed
    (+ learned guess) timed as consecutive runs, and their errors after the checkpoint steps."""
    sh = dict(s, RIEMANN="hlle")
    sp = dict(sh, PRECISION="fp64", C2P_GUESS="plain")
    sm = dict(sh, PRECISION="mixed", C2P_GUESS="learned" if mlp is not None else "plain")
    zu = s["NX"] * s["NY"] * s["NZ"] * nsteps
    log(f"[bench] HLLE companion: {nsteps} consecutive steps, plain and mixed")
    p = run_timed(sp, pr0, nsteps, 0, repeat=False)
    m = run_timed(sm, pr0, nsteps, 0, mlp=mlp, repeat=False)
    log(f"[bench]   plain {p['elapsed'] * 1e3 / nsteps:.2f} ms/step, mixed {m['elapsed'] * 1e3 / nsteps:.2f} ms/step")
    ev = accuracy(sp, sm, pr0, set(checkpoints), mlp)
    return {"workload": f"{DEFAULT_WORKLOAD} with RIEMANN=hlle, {nsteps} consecutive steps",
            "fp64_plain": timed_entry(p, zu, nsteps), "mixed": timed_entry(m, zu, nsteps),
            "speedup_vs_fp64_plain": round(p["elapsed"] / m["elapsed"], 3), "errors_vs_fp64_plain": ev}


def config_bench(name, steps, warmup, checkpoints, mlp, flips=False):
    """Another benchmark config on one GPU: timed plain FP64 and mixed steps (every step from the
    seeded state), each run's dominant-kernel roofline, and the mixed run's errors against the plain
    one after the checkpoint steps."""
    from astrofd_amd import ic
    s = load_workload(name)
    t0 = time.perf_counter()
    pr0 = ic.initial_state(s)
    log(f"[bench] {name}: initial state {pr0.shape} in {time.perf_counter() - t0:.1f}s")
    rm = s["PHYSICS"] == "rmhd"
    use_mlp = mlp if rm else None
    sp = dict(s, PRECISION="fp64", C2P_GUESS="plain")
    sm = dict(s, PRECISION="mixed", C2P_GUESS="learned" if use_mlp is not None else "plain")
    zu = s["NX"] * s["NY"] * s["NZ"] * steps
    p = run_timed(sp, pr0, steps, warmup)
    log(f"[bench]   {name} plain {p['elapsed'] * 1e3 / steps:.2f} ms/step")
    m = run_timed(sm, pr0, steps, warmup, mlp=use_mlp)
    log(f"[bench]   {name} mixed {m['elapsed'] * 1e3 / steps:.2f} ms/step")
    out = {"workload": f"{name}: {workload_desc(s)}", "steps": steps, "warmup": warmup,
           "fp64_plain": dict(timed_entry(p, zu, steps), roofline=kernel_roofline(s, p["kernels"], False, False, name)),
           "mixed": dict(timed_entry(m, zu, steps), c2p_guess=sm["C2P_GUESS"],
                         roofline=kernel_roofline(s, m["kernels"], True, use_mlp is not None, name)),
           "speedup_vs_fp64_plain": round(p["elapsed"] / m["elapsed"], 3)}
    out["errors_vs_fp64_plain"] = accuracy(sp, sm, pr0, set(checkpoints), use_mlp, flips=flips)
    return out


def workload_desc(s):
    rm = s["PHYSICS"] == "rmhd"
    recon = {"weno": "WENO5", "ppm": "PPM"}.get(str(s["RECON"]).lower(), f"MUSCL({s['LIMITER']})")
    riem = "HLLD(hll_d_flux_dir)" if rm and s["RIEMANN"] == "hlld_full" else str(s["RIEMANN"]).upper()
    extra = ""
    if s.get("RESISTIVE_ENABLED"):
        extra += f", resistivity {s['RESISTIVITY']:g}"
    if s.get("TWO_TEMPERATURE"):
        extra += f", Te/Ti (TEI_TAU {s['TEI_TAU']:g})"
    return (f"{'RMHD' if rm else 'SRHD'} {s['NX']}^3 {recon} + {riem} + SSPRK{s['RK_ORDER']}, "
            f"Gamma={s['GAMMA']:.4g}, CFL {s['CFL']}{extra}")


def kernel_ms(k, timing_steps):
    return {"face_x": round(k["ms_face"][0] / timing_steps, 3), "face_y": round(k["ms_face"][1] / timing_steps, 3),
            "face_z": round(k["ms_face"][2] / timing_steps, 3),
            "update_c2p": round(k["ms_update"] / timing_steps, 3), "other": round(k["ms_other"] / timing_steps, 3)}


def c2p_report(st):
    cells = max(st["c2p_cells"], 1)
    return {"hydro_iters_per_cell": round(st["hydro_iters"] / cells, 3),
            "newton_evals_per_cell": round(st["newton_evals"] / cells, 3),
            "learned_accept_frac": round(st["learned_ok"] / cells, 5),
            "bisection_frac": round(st["bisections"] / cells, 5), "invalid_frac": round(st["invalid"] / cells, 5)}


def copy_probe(gib):
    """Plain device kernels of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters: a copy
    (reads and writes n x 8 B), a sum (reads n x 8 B) and a fill (writes n x 8 B), three launches each."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    s = None
    for _ in range(3):
        s = a.sum()
    for _ in range(3):
        b.fill_(2.0)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: {n * 8} bytes read + {n * 8} bytes written per copy; sum of {n * 8} bytes "
        f"({float(s):.0f}); fill of {n * 8} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 plain run")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the lockstep accuracy pass")
    ap.add_argument("--no-sources", action="store_true", help="skip the post-step source kernel measurement")
    ap.add_argument("--sources-only", action="store_true",
                    help="measure only the post-step source kernel and print its JSON line (profiling)")
    ap.add_argument("--accuracy-steps", default="1,2,5,10")
    ap.add_argument("--face-path-ab", action="store_true",
                    help="also time the mixed run with one thread per face (afd_params.face_path = 1)")
    ap.add_argument("--face-path-sweep", action="store_true",
                    help="time the mixed run with every face-kernel variant (afd_params.face_path 1..4)")
    ap.add_argument("--evolve", action="store_true",
                    help="time consecutive steps of one run instead of steps from the seeded state")
    ap.add_argument("--evolve-steps", type=int, default=10,
                    help="consecutive steps timed beside the headline (0 = skip)")
    ap.add_argument("--copy-probe", type=float, default=0.0, help="GiB of a plain device copy (profiling)")
    ap.add_argument("--no-hlle", action="store_true", help="skip the 50-step HLLE companion")
    ap.add_argument("--no-configs", action="store_true", help="skip the SRHD 256^3 and RMHD 512^3 entries")
    ap.add_argument("--no-network", action="store_true", help="skip the MLP rate measurement")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): the step is built and measured for one "
              "MI355X; sharding across GPUs is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from astrofd_amd import ic
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
    s = load_workload(a.workload)
    log(f"[bench] workload {a.workload}: {s['NX']}x{s['NY']}x{s['NZ']} {s['PHYSICS']} {s['RECON']}/{s['RIEMANN']} "
        f"rk{s['RK_ORDER']} steps={a.steps} warmup={a.warmup}")
    t0 = time.perf_counter()
    pr0 = ic.initial_state(s)
    log(f"[bench] initial state {pr0.shape} in {time.perf_counter() - t0:.1f}s")
    zu = s["NX"] * s["NY"] * s["NZ"] * a.steps
    if a.sources_only:
        print(json.dumps({"workload": a.workload, "post_step_sources": sources_roofline(s, pr0)}), flush=True)
        return

    mlp = None
    if s["PHYSICS"] == "rmhd" and os.path.exists(MLP_WEIGHTS):
        with np.load(MLP_WEIGHTS, allow_pickle=False) as f:
            mlp = np.asarray(f["w"], dtype=np.float32)
    sp = dict(s, PRECISION="fp64", C2P_GUESS="plain")
    sm = dict(s, PRECISION="mixed", C2P_GUESS="learned" if mlp is not None else "plain")

    plain = None
    if not a.no_plain:
        log("[bench] plain FP64 run")
        plain = run_timed(sp, pr0, a.steps, a.warmup, repeat=not a.evolve)
        log(f"[bench]   {plain['elapsed'] * 1e3 / a.steps:.2f} ms/step, {zu / plain['elapsed']:.4e} ZU/s")
    log(f"[bench] mixed run (c2p guess: {sm['C2P_GUESS']})")
    mixed = run_timed(sm, pr0, a.steps, a.warmup, mlp=mlp, repeat=not a.evolve)
    log(f"[bench]   {mixed['elapsed'] * 1e3 / a.steps:.2f} ms/step, {zu / mixed['elapsed']:.4e} ZU/s")

    mixed_plain = None
    if mlp is not None:
        log("[bench] mixed run with the plain con2prim (A/B of the learned guess)")
        mixed_plain = run_timed(dict(sm, C2P_GUESS="plain"), pr0, a.steps, a.warmup, repeat=not a.evolve)
        log(f"[bench]   {mixed_plain['elapsed'] * 1e3 / a.steps:.2f} ms/step")

    mixed_prev = None
    if s["PHYSICS"] == "rmhd":
        log("[bench] mixed run with the previous-state Newton start")
        mixed_prev = run_timed(dict(sm, C2P_GUESS="previous"), pr0, a.steps, a.warmup, repeat=not a.evolve)
        log(f"[bench]   {mixed_prev['elapsed'] * 1e3 / a.steps:.2f} ms/step")

    mixed_pf = None
    if a.face_path_ab:
        log("[bench] mixed run with one thread per face (face_path = 1)")
        mixed_pf = run_timed(sm, pr0, a.steps, a.warmup, mlp=mlp, repeat=not a.evolve, face_path=1)
        log(f"[bench]   {mixed_pf['elapsed'] * 1e3 / a.steps:.2f} ms/step")

    if a.face_path_sweep:
        sweep = {}
        for fp in (1, 2, 3, 4):
            log(f"[bench] mixed run with face_path = {fp}")
            r = run_timed(sm, pr0, a.steps, a.warmup, mlp=mlp, repeat=not a.evolve, face_path=fp)
            sweep[str(fp)] = {"ms_per_step": round(r["elapsed"] * 1e3 / a.steps, 3),
                              "kernel_ms_per_step": kernel_ms(r["kernels"], a.steps)}
            log(f"[bench]   {sweep[str(fp)]}")

    value = zu / mixed["elapsed"]
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    rm = s["PHYSICS"] == "rmhd"
    desc = workload_desc(s)
    res = {
        "metric": ("zone-updates/s RMHD 256³ WENO5+HLLD: mixed+learned vs FP64 plain at L1≤1e-6"
                   if a.workload == DEFAULT_WORKLOAD else f"zone-updates/s {desc}: mixed vs FP64 plain"),
        "value": round(value, 1), "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(mixed["elapsed"] * 1e3 / a.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64",
        "data": ("synthetic (seeded IC: 1% density noise, Bx=0.1, sigma=0.1 toroidal nozzle, curl-A dB rms 1e-3)" if rm
                 else "synthetic (seeded IC: 1% density noise, W = 7.09 top-hat nozzle)"),
        "config": {"workload": f"{a.workload}: {desc}, one step = fills + SSPRK{s['RK_ORDER']} + fills, "
                               + ("consecutive steps" if a.evolve else "every timed step from the seeded state"),
                   "grid": [s["NX"], s["NY"], s["NZ"]], "ng": s["NG"], "precision": "mixed",
                   "c2p_guess": sm["C2P_GUESS"]},
    }
    res["roofline"] = kernel_roofline(s, mixed["kernels"], True, mlp is not None, a.workload)
    res["kernel_ms_per_step"] = kernel_ms(mixed["kernels"], a.steps)
    step_bytes = (13 if s["RK_ORDER"] == 3 else 8) * nvars(s) * 8
    res["step_roofline"] = {"bytes_per_zone_update": step_bytes, "achieved": round(step_bytes * value / 1e9, 2),
                            "unit": "GB/s", "frac": round(step_bytes * value / 1e9 / HBM_PEAK_GBS, 5)}
    res["c2p"] = c2p_report(mixed["stats"])
    if mixed_plain is not None:
        res["mixed_plain_c2p"] = {"value": round(zu / mixed_plain["elapsed"], 1),
                                  "ms_per_step": round(mixed_plain["elapsed"] * 1e3 / a.steps, 3),
                                  "kernel_ms_per_step": kernel_ms(mixed_plain["kernels"], a.steps),
                                  "c2p": c2p_report(mixed_plain["stats"])}
    if mixed_prev is not None:
        res["mixed_previous_guess"] = {"value": round(zu / mixed_prev["elapsed"], 1),
                                       "ms_per_step": round(mixed_prev["elapsed"] * 1e3 / a.steps, 3),
                                       "kernel_ms_per_step": kernel_ms(mixed_prev["kernels"], a.steps),
                                       "c2p": c2p_report(mixed_prev["stats"])}
    if a.face_path_sweep:
        res["face_path_sweep"] = sweep
    if mixed_pf is not None:
        res["mixed_per_face_kernels"] = {"value": round(zu / mixed_pf["elapsed"], 1),
                                         "ms_per_step": round(mixed_pf["elapsed"] * 1e3 / a.steps, 3),
                                         "kernel_ms_per_step": kernel_ms(mixed_pf["kernels"], a.steps)}
    if plain is not None:
        res["fp64_plain"] = {"value": round(zu / plain["elapsed"], 1),
                             "ms_per_step": round(plain["elapsed"] * 1e3 / a.steps, 3),
                             "kernel_ms_per_step": kernel_ms(plain["kernels"], a.steps),
                             "roofline": kernel_roofline(s, plain["kernels"], False, False, a.workload),
                             "c2p": c2p_report(plain["stats"])}
        res["speedup_vs_fp64_plain"] = round(plain["elapsed"] / mixed["elapsed"], 3)
        if not a.no_accuracy:
            cps = sorted({int(x) for x in a.accuracy_steps.split(",") if x.strip()} | {1})
            log(f"[bench] accuracy pass (steps {cps})")
            ev = accuracy(sp, sm, pr0, set(cps), mlp)
            # the headline: one step from the seeded state, as timed; later steps of the same run
            # show how the reference scheme's own divergence (non-finite counts) takes over
            res["errors_vs_fp64_plain"] = ev[0]
            res["errors_evolution"] = ev
            res["accuracy_scope"] = ("the metric's L1(rho) <= 1e-6 is for one step from the seeded state, the timed "
                                     "unit; errors_evolution and hlle_50 give later steps beside the FP64 run's own "
                                     "response to 1-ulp input noise (fp64_ulp_spread)")
            res["fp64_ulp_spread"] = dict(ev[0]["fp64_ulp_spread"], step=1)
        if a.evolve_steps > 0 and not a.evolve:
            zu_e = s["NX"] * s["NY"] * s["NZ"] * a.evolve_steps
            log(f"[bench] {a.evolve_steps} consecutive steps (plain, mixed)")
            ep = run_timed(sp, pr0, a.evolve_steps, 0, repeat=False)
            em = run_timed(sm, pr0, a.evolve_steps, 0, mlp=mlp, repeat=False)
            res["consecutive_steps"] = {"steps": a.evolve_steps, "fp64_plain": round(zu_e / ep["elapsed"], 1),
                                        "mixed": round(zu_e / em["elapsed"], 1),
                                        "mixed_c2p": c2p_report(em["stats"])}
    if rm and not a.no_hlle and a.workload == DEFAULT_WORKLOAD:
        res["hlle_50"] = hlle_leg(s, pr0, mlp)
    if mlp is not None and not a.no_network:
        log("[bench] MLP forward rate")
        res["network"] = network_rate(mlp)
    if not a.no_sources:
        log("[bench] post-step sources kernel")
        res["post_step_sources"] = sources_roofline(s, pr0)
    if not a.no_configs and a.workload == DEFAULT_WORKLOAD:
        del pr0
        res["srhd_jet_256"] = config_bench("srhd_jet_256", 10, 2, (1, 10, 100), None, flips=True)
        res["rmhd_jet_512"] = config_bench("rmhd_jet_512", 3, 1, (1,), mlp)
    if not a.no_cpu_baseline:
        log("[bench] CPU oracle sample")
        res["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
