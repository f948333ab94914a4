#!/usr/bin/env python3
"""Benchmark of the SRHD/RMHD finite-volume step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmhd_jet_256]

Workload (default): the RMHD 256^3 jet, WENO5 + HLLD + SSPRK3, Gamma = 4/3
(astrofd_amd/config/rmhd_jet_256.json).  A step is one pass of the hot path over
the whole grid, as the reference driver runs it: ghost/nozzle fills, one SSPRK3
step, ghost/nozzle fills.  Inputs are resident in HBM before the timed region.

Two runs from the same seeded initial state:
  plain  FP64 everything, plain con2prim (the baseline of the metric)
  mixed  FP32 reconstruction + Riemann fluxes, FP64 Lorentz factor / con2prim /
         SSPRK accumulation, learned con2prim guess when weights are available
`value` is the mixed run's zone-updates/s; its relative L1 error in rho, p and W
against the plain run after the same steps is reported beside it.

Prints ONE JSON line.  Also: the dominant kernel's roofline from HIP events over
a timed pass, and the CPU oracle's rate on a bounded sample (64^3, one step) on
this host's cores.
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
    """Relative L1 and max |diff| of rho, p, W over interior cells (b is the reference)."""
    sl = (slice(ng, -ng),) * 3
    out = {}
    for name, fa, fb in (("rho", a[0][sl], b[0][sl]), ("p", a[4][sl], b[4][sl]),
                         ("W", lorentz(a)[sl], lorentz(b)[sl])):
        d = np.abs(fa - fb)
        out[name] = {"L1_rel": float(d.sum() / np.abs(fb).sum()), "max_abs": float(d.max())}
    return out


def sync(d):
    d.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def run_path(s, pr0, steps, warmup, mlp=None, timing_steps=0):
    from astrofd_amd.domain import Domain
    d = Domain.from_settings(s)
    try:
        d.set_prim(pr0)
        d.set_nozzle(s)
        if mlp is not None:
            d.set_mlp_weights(mlp)
        rng = np.random.default_rng(1234)
        dx = min(s["Lx"] / s["NX"], s["Ly"] / s["NY"], s["Lz"] / s["NZ"])
        d.apply_bcs(rng)

        def one_step():
            amax = d.max_char_speed() if s["PHYSICS"] == "hydro" else 1.0
            dt = s["CFL"] * dx / max(amax, 1e-12)
            d.apply_bcs(rng)
            d.step(dt)
            d.apply_bcs(rng)
            d.two_temperature(dt)

        for _ in range(warmup):
            one_step()
        d.set_instrumentation(stats=True, timing=False)
        d.reset_stats()
        sync(d)
        t0 = time.perf_counter()
        for _ in range(steps):
            one_step()
        sync(d)
        elapsed = time.perf_counter() - t0
        st = d.stats()
        kern = None
        if timing_steps:
            d.set_instrumentation(stats=False, timing=True)
            d.reset_stats()
            for _ in range(timing_steps):
                one_step()
            kern = d.stats()
            d.set_instrumentation(stats=False, timing=False)
        out = d.get_prim()
        return dict(elapsed=elapsed, stats=st, kernels=kern, state=out, nonfinite=int(np.count_nonzero(~np.isfinite(out))))
    finally:
        d.close()


def kernel_roofline(s, kern, mixed):
    """Dominant kernel and its achieved algorithmic HBM bandwidth (bytes per launch / avg duration).

    Algorithmic bytes per launch (compulsory traffic of one launch over its units):
      face kernel, one thread per face: read the nb primitives of the face's cell (8 B each) and
        write the face flux (nb x flux size; twice for the RMHD PPM/WENO x faces, plus the two carried
        speeds read there and written by the z kernel).
      stage update, one thread per cell: read U0 (stage > 0), U_s-1 (stage > 0), the stage input
        primitives' GLM component and the three face fluxes of the cell's left faces (nb each, the right
        faces are the neighbours' left faces), write U_s and the stage output primitives.
    """
    nb = 9 if s["PHYSICS"] == "rmhd" else 5
    fb = 4 if mixed else 8
    NX, NY, NZ = s["NX"], s["NY"], s["NZ"]
    ncells = NX * NY * NZ
    full = (NX + 2 * s["NG"]) * (NY + 2 * s["NG"]) * (NZ + 2 * s["NG"])
    carried = nb == 9 and s["RECON"] != "muscl"
    per_face = [nb * 8 + nb * fb, nb * 8 + nb * fb, nb * 8 + nb * fb]
    if carried:
        per_face[0] += nb * fb + 2 * fb
        per_face[2] += 2 * fb
    faces = [(NX + 1) * NY * NZ, NX * (NY + 1) * NZ, NX * NY * (NZ + 1)]
    # update: 3 stages average of U0 + Uc reads (2/3 of stages), writes Uc + P, flux reads 3 nb
    upd_per_cell = nb * 8 * (2 * 2 / 3) + nb * 8 * 2 + 3 * nb * fb
    cands = []
    for dkey, name in ((0, "face_x"), (1, "face_y"), (2, "face_z")):
        n = kern["launches_face"][dkey]
        if n:
            cands.append((kern["ms_face"][dkey] / n, name, per_face[dkey] * faces[dkey]))
    if kern["launches_update"]:
        cands.append((kern["ms_update"] / kern["launches_update"], "stage_update_c2p", upd_per_cell * full))
    ms, name, nbytes = max(cands)
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_ms": round(ms, 4),
            "bytes_per_launch": int(nbytes)}


def cpu_baseline(s, n=64):
    """The CPU oracle (C restatement of the reference, OpenMP) on an n^3 sample of the same workload:
    one step (fills + SSPRK step + fills), all host cores."""
    import oracle as O
    from astrofd_amd import ic
    t = dict(s)
    t.update(NX=n, NY=n, NZ=n)
    pr = ic.initial_state(t)
    rm = t["PHYSICS"] == "rmhd"
    P = O.make_params("rmhd" if rm else "hydro", pr.shape[1:], nv=pr.shape[0], ng=t["NG"],
                      gamma=t["GAMMA"] if rm else None, p_max=t["P_MAX"], v_max=t["V_MAX"], glm_ch=t["GLM_CH"],
                      glm_cp=t["GLM_CP"], limiter=t["LIMITER"], recon=t["RECON"], riemann=t["RIEMANN"],
                      rk_order=t["RK_ORDER"], n_passive=t["N_PASSIVE"])
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    O.set_num_threads(cores)
    dx = 1.0 / n
    bamb = tuple(t.get("B_AMB") or (0.0, 0.0, 0.0))

    def bcs(a):
        O.apply_outflow_yz(P, a, t["BC_Y"] == "outflow", t["BC_Z"] == "outflow")
        O.apply_outflow_right_x(P, a)
        O.apply_nozzle_left_x(P, a, dx, dx, dx, n, n, t["JET_CENTER"][1], t["JET_CENTER"][2], t, None, bamb)

    bcs(pr)
    t0 = time.perf_counter()
    bcs(pr)
    amax = O.max_char_speed(P, pr) if not rm else 1.0
    pr = O.step(P, pr, dx, dx, dx, t["CFL"] * dx / amax)
    bcs(pr)
    el = time.perf_counter() - t0
    return {"value": round(n ** 3 / el, 1), "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"one step of the same workload at {n}^3 (C oracle, OpenMP, {cores} threads, "
                      f"{el:.2f} s wall)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmhd_jet_256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 plain run")
    ap.add_argument("--timing-steps", type=int, default=2)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): the step is built and measured for one "
              "MI355X; multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from astrofd_amd import ic
    s = load_workload(a.workload)
    log(f"[bench] workload {a.workload}: {s['NX']}x{s['NY']}x{s['NZ']} {s['PHYSICS']} {s['RECON']}/{s['RIEMANN']} "
        f"rk{s['RK_ORDER']} steps={a.steps} warmup={a.warmup}")
    t0 = time.perf_counter()
    pr0 = ic.initial_state(s)
    log(f"[bench] initial state {pr0.shape} in {time.perf_counter() - t0:.1f}s")
    zu = s["NX"] * s["NY"] * s["NZ"] * a.steps

    plain = None
    if not a.no_plain:
        sp = dict(s, PRECISION="fp64", C2P_GUESS="plain")
        log("[bench] plain FP64 run")
        plain = run_path(sp, pr0, a.steps, a.warmup, timing_steps=a.timing_steps)
        log(f"[bench]   {plain['elapsed'] * 1e3 / a.steps:.2f} ms/step, {zu / plain['elapsed']:.4e} ZU/s")

    mlp = None
    if s["PHYSICS"] == "rmhd" and os.path.exists(MLP_WEIGHTS):
        mlp = np.load(MLP_WEIGHTS)["w"].astype(np.float32)
    sm = dict(s, PRECISION="mixed", C2P_GUESS="learned" if mlp is not None else "plain")
    log(f"[bench] mixed run (c2p guess: {sm['C2P_GUESS']})")
    mixed = run_path(sm, pr0, a.steps, a.warmup, mlp=mlp, timing_steps=a.timing_steps)
    log(f"[bench]   {mixed['elapsed'] * 1e3 / a.steps:.2f} ms/step, {zu / mixed['elapsed']:.4e} ZU/s")

    value = zu / mixed["elapsed"]
    res = {
        "metric": "zone-updates/s RMHD 256³ WENO5+HLLD: mixed+learned vs FP64 plain at L1≤1e-6",
        "value": round(value, 1), "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(mixed["elapsed"] * 1e3 / a.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded IC: 1% density noise, Bx=0.1, sigma=0.1 toroidal nozzle, curl-A dB rms 1e-3)",
        "config": {"workload": f"{a.workload}: RMHD {s['NX']}^3 WENO5 + HLLD(hll_d_flux_dir) + SSPRK3, Gamma=4/3, "
                               f"CFL {s['CFL']}, one step = fills + SSPRK3 + fills",
                   "grid": [s["NX"], s["NY"], s["NZ"]], "ng": s["NG"], "precision": "mixed",
                   "c2p_guess": sm["C2P_GUESS"]},
    }
    if plain is not None:
        res["fp64_plain"] = {"value": round(zu / plain["elapsed"], 1),
                             "ms_per_step": round(plain["elapsed"] * 1e3 / a.steps, 3)}
        res["speedup_vs_fp64_plain"] = round(plain["elapsed"] / mixed["elapsed"], 3)
        res["errors_vs_fp64_plain"] = errors(mixed["state"], plain["state"], s["NG"])
        res["nonfinite"] = {"plain": plain["nonfinite"], "mixed": mixed["nonfinite"]}
    st = mixed["stats"]
    cells = max(st["c2p_cells"], 1)
    res["c2p"] = {"hydro_iters_per_cell": round(st["hydro_iters"] / cells, 3),
                  "newton_evals_per_cell": round(st["newton_evals"] / cells, 3),
                  "learned_accept_frac": round(st["learned_ok"] / cells, 5),
                  "bisections": st["bisections"], "invalid": st["invalid"]}
    if plain is not None:
        sp_ = plain["stats"]
        cp = max(sp_["c2p_cells"], 1)
        res["c2p"]["plain_hydro_iters_per_cell"] = round(sp_["hydro_iters"] / cp, 3)
        res["c2p"]["plain_newton_evals_per_cell"] = round(sp_["newton_evals"] / cp, 3)
    if mixed["kernels"]:
        res["roofline"] = kernel_roofline(s, mixed["kernels"], True)
        k = mixed["kernels"]
        res["kernel_ms_per_step"] = {
            "face_x": round(k["ms_face"][0] / a.timing_steps, 3), "face_y": round(k["ms_face"][1] / a.timing_steps, 3),
            "face_z": round(k["ms_face"][2] / a.timing_steps, 3),
            "update_c2p": round(k["ms_update"] / a.timing_steps, 3), "other": round(k["ms_other"] / a.timing_steps, 3)}
        nb = 9 if s["PHYSICS"] == "rmhd" else 5
        step_bytes = (13 if s["RK_ORDER"] == 3 else 8) * nb * 8
        res["step_roofline"] = {"bytes_per_zone_update": step_bytes,
                                "achieved": round(step_bytes * value / 1e9, 2), "unit": "GB/s",
                                "frac": round(step_bytes * value / 1e9 / HBM_PEAK_GBS, 5)}
    if plain is not None and plain["kernels"]:
        k = plain["kernels"]
        res["fp64_plain"]["kernel_ms_per_step"] = {
            "face_x": round(k["ms_face"][0] / a.timing_steps, 3), "face_y": round(k["ms_face"][1] / a.timing_steps, 3),
            "face_z": round(k["ms_face"][2] / a.timing_steps, 3),
            "update_c2p": round(k["ms_update"] / a.timing_steps, 3), "other": round(k["ms_other"] / a.timing_steps, 3)}
        res["fp64_plain"]["roofline"] = kernel_roofline(s, k, False)
    if not a.no_cpu_baseline:
        log("[bench] CPU oracle sample")
        res["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
