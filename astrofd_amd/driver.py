"""Jet driver on one MI355X, in the shape of solvers/srhd3d_mpi_muscl.py.

    python -m astrofd_amd.driver --config astrofd_amd/config/srhd_jet_32.json [--nx N ...]

Same settings keys and CLI (utils/settings.py:100-107, plus --max-steps), the same loop
(solvers/srhd3d_mpi_muscl.py:276-330): CFL time step (SRHD: max characteristic speed; RMHD:
amax = 1), ghost/nozzle fills, one SSPRK2/3 step, fills again, post-step operators (Israel-Stewart
dissipation; cooling/heating, two-temperature and ion chemistry fused into one pass), progress
print, NaN check, npz outputs and checkpoints with the reference's keys and meta, and the
diagnostics CSVs of utils/diagnostics.py (max Gamma and inlet flux, centerline, div B), whose
reductions run on the GPU.  The whole
domain is one GPU-resident block, so the MPI slab exchange (:62-126) has no counterpart: the
fills provide every ghost zone.  The state stays in HBM between steps; it is copied to the
host only for outputs, checkpoints and the NaN check.
"""
import glob
import os
import sys
import time
from datetime import datetime

import numpy as np

from . import ic
from .domain import Domain
from .settings import load_settings

SMALL = 1e-12


def make_run_dir(base="results", unique=False):
    """results/YYYY-MM-DD/ (or .../HH-MM-SS/ with RESULTS_UNIQUE), as utils/io_utils.py:5-18."""
    now = datetime.now()
    path = os.path.join(base, now.strftime("%Y-%m-%d"))
    if unique:
        path = os.path.join(path, now.strftime("%H-%M-%S"))
    os.makedirs(path, exist_ok=True)
    return path


def latest_run_dir(base="results"):
    runs = sorted(d for d in glob.glob(os.path.join(base, "*")) if os.path.isdir(d))
    return runs[-1] if runs else None


def latest_checkpoint_in_dir(run_dir, rank=0):
    files = sorted(glob.glob(os.path.join(run_dir, f"checkpoint_rank{rank:04d}_step*.npz")))
    return files[-1] if files else None


def resolve_restart_path(restart_path, rank=0):
    """solvers/srhd3d_mpi_muscl.py:176-191: None, "latest", a run directory or a file."""
    if restart_path is None:
        return None
    if isinstance(restart_path, str) and restart_path.lower() in ("none", ""):
        return None
    if isinstance(restart_path, str) and restart_path.lower() == "latest":
        run_dir = latest_run_dir()
        return latest_checkpoint_in_dir(run_dir, rank) if run_dir else None
    if os.path.isdir(restart_path):
        return latest_checkpoint_in_dir(restart_path, rank)
    if os.path.isfile(restart_path):
        return restart_path
    return None


def load_checkpoint(path):
    """(prim, t, step) of a checkpoint npz (meta = [dx, dy, dz, t, step])."""
    with np.load(path, allow_pickle=False) as data:
        pr = np.array(data["prim"])
        meta = np.array(data["meta"]) if "meta" in data.files else None
    if meta is None or len(meta) < 5:
        return pr, 0.0, 0
    return pr, float(meta[3]), int(meta[4])


def output_fields(s, pr):
    """The arrays the reference writes per output (solvers/srhd3d_mpi_muscl.py:349-393)."""
    f = {}
    if int(s.get("N_TRACERS", 0)) > 0:
        names = s.get("TRACER_NAMES") or []
        for t in range(int(s["N_TRACERS"])):
            f[names[t] if t < len(names) else f"tracer{t}"] = pr[int(s["TRACER_OFFSET"]) + t]
        f["tracer_names"] = np.array(names, dtype=np.str_)
    if int(s.get("N_THERMO", 0)) > 0:
        f["Te"] = pr[int(s["THERMO_OFFSET"])]
        f["Ti"] = pr[int(s["THERMO_OFFSET"]) + 1]
    if int(s.get("N_CHEM", 0)) > 0:
        names = s.get("CHEM_NAMES") or []
        for m in range(int(s["N_CHEM"])):
            f[names[m] if m < len(names) else f"chem{m}"] = pr[int(s["CHEM_OFFSET"]) + m]
    base = dict(rho=pr[0], vx=pr[1], vy=pr[2], vz=pr[3], p=pr[4])
    if s["PHYSICS"] in ("rmhd", "grmhd"):
        base.update(Bx=pr[5], By=pr[6], Bz=pr[7], psi=pr[8])
    elif int(s.get("TRACER_OFFSET", 5)) > 5:  # the dissipative fields (:362-372)
        for m, name in enumerate(("pi", "pixx", "piyy", "pizz", "pixy", "pixz", "piyz", "qx", "qy", "qz")):
            base[name] = pr[5 + m]
    base.update(f)
    return base


def _append_csv(path, header, row):
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(header + "\n")
        f.write(row + "\n")


def write_diagnostics(d, s, run_dir, step, t, dt, amax, dx):
    """The reference's diagnostics.csv, centerline.csv and (RMHD) divb.csv rows
    (utils/diagnostics.py:46-54, :111-126, :177-181), from the device reductions."""
    g = d.diagnostics()
    flux = g["inlet_flux"]
    _append_csv(os.path.join(run_dir, "diagnostics.csv"), "step,time,dt,amax,maxGamma,inletFlux_signed,inletFlux_abs",
                f"{step},{t:.8e},{dt:.8e},{amax:.6e},{g['max_gamma']:.6e},{flux:.6e},{abs(flux):.6e}")
    cl = d.centerline()
    path = os.path.join(run_dir, "centerline.csv")
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write("step,time,i,x,Gamma,rho,p,vx\n")
        for i in range(cl.shape[1]):
            f.write(f"{step},{t:.8e},{i},{(i + 0.5) * dx:.8e},{cl[0, i]:.8e},{cl[1, i]:.8e},{cl[2, i]:.8e},"
                    f"{cl[3, i]:.8e}\n")
    if s["PHYSICS"] in ("rmhd", "grmhd"):
        _append_csv(os.path.join(run_dir, "divb.csv"), "step,time,divB_max,divB_rms,divB_mean",
                    f"{step},{t:.8e},{g['divb_max']:.6e},{g['divb_rms']:.6e},{g['divb_mean']:.6e}")
    return g


def load_mlp(s):
    from . import learned
    path = s.get("C2P_MODEL") or learned.MODEL_PATH
    return learned.load(path)


def run(s, log=print):
    """The time loop; returns (final primitives, t, steps)."""
    NX, NY, NZ, NG = s["NX"], s["NY"], s["NZ"], s["NG"]
    dx, dy, dz = s["Lx"] / NX, s["Ly"] / NY, s["Lz"] / NZ
    restart_path = resolve_restart_path(s.get("RESTART_PATH"))
    run_dir = os.path.dirname(restart_path) if restart_path else \
        make_run_dir(base="results", unique=s.get("RESULTS_UNIQUE", False))
    log(f"[startup] run directory: {run_dir}")
    log(f"[startup] ranks=1 grid={NX}x{NY}x{NZ} NG={NG} debug={s['DEBUG']} device=MI355X "
        f"precision={s['PRECISION']} c2p_guess={s['C2P_GUESS']}")
    log("[startup] params:"
        f" P_EQ={s['P_EQ']:.6g}, P_AMB={s['P_AMB']:.6g}, RHO_AMB={s['RHO_AMB']:.6g},"
        f" GAMMA_JET={s['GAMMA_JET']:.6g}, ETA_RHO={s['ETA_RHO']:.6g},"
        f" JET_RADIUS={s['JET_RADIUS']:.3f}, NOZZLE_TURB={s['NOZZLE_TURB']}")
    log(f"[startup] physics={s['PHYSICS']}  B_INIT={s['B_INIT']}  B0={s['B0']}")

    rng = np.random.default_rng(1234)  # rank 0's generator (solvers/srhd3d_mpi_muscl.py:230)
    t, step = 0.0, 0
    if restart_path:
        pr, t, step = load_checkpoint(restart_path)
        log(f"[restart] loaded {restart_path} at t={t:.6e} step={step}")
    else:
        pr = ic.initial_state(s)

    d = Domain.from_settings(s)
    try:
        d.set_prim(pr)
        d.set_nozzle(s)
        if s["PHYSICS"] in ("rmhd", "grmhd") and s["C2P_GUESS"] == "learned":
            d.set_mlp_weights(load_mlp(s))
        if not restart_path:
            d.apply_bcs(rng)
        if s["DEBUG"]:
            d.compute_rhs()
            log("[jit] RHS kernel compiled and first call done.")

        T_END, max_steps = float(s["T_END"]), int(s.get("MAX_STEPS", 0) or 0)
        hydro = s["PHYSICS"] in ("hydro", "grhd")
        t0 = time.perf_counter()
        while t < T_END and (max_steps <= 0 or step < max_steps):
            amax = d.max_char_speed() if hydro else 1.0
            dt = s["CFL"] * min(dx, dy, dz) / max(amax, SMALL)
            if t + dt > T_END:
                dt = T_END - t
            d.apply_bcs(rng)
            d.step(dt)
            d.apply_bcs(rng)
            d.post_step_sources(dt)  # dissipation, cooling/heating, two-temperature, chemistry (:323-326)
            t += dt
            step += 1
            if step % int(s["PRINT_EVERY"]) == 0 or abs(t - T_END) < 1e-14:
                log(f"[rank0] t={t:.5f} dt={dt:.3e} amax={amax:.3f} step={step}")
            cne = int(s.get("CHECK_NAN_EVERY", 0) or 0)
            if cne > 0 and step % cne == 0 and d.count_nonfinite() > 0:
                log("[warn] NaN/Inf detected in primitives!")
                if s.get("ASSERTS", False):
                    raise RuntimeError("NaN/Inf in primitives")
            if step % int(s["OUT_EVERY"]) == 0 or abs(t - T_END) < 1e-14:
                pr = d.get_prim()
                fname = os.path.join(run_dir, f"jet3d_rank0000_step{step:06d}.npz")
                kind = "3D RMHD/GRMHD block (with ghosts)." if s["PHYSICS"] in ("rmhd", "grmhd") else \
                    "3D SRHD/GRHD block (with ghosts)."
                np.savez(fname, **output_fields(s, pr), meta=np.array([dx, dy, dz, t], dtype=np.float64),
                         comment=kind)
                log(f"[io] wrote {fname}")
                g = write_diagnostics(d, s, run_dir, step, t, dt, amax, dx)
                log(f"[diag] step={step} maxGamma={g['max_gamma']:.3f} "
                    f"inletFlux_signed={g['inlet_flux']:.3e} inletFlux_abs={abs(g['inlet_flux']):.3e}")
            ce = int(s.get("CHECKPOINT_EVERY", 0) or 0)
            if ce > 0 and step % ce == 0:
                ckpt = os.path.join(run_dir, f"checkpoint_rank0000_step{step:06d}.npz")
                np.savez(ckpt, prim=d.get_prim(), meta=np.array([dx, dy, dz, t, step], dtype=np.float64),
                         comment="Checkpoint (with ghosts).")
                log(f"[ckpt] wrote {ckpt}")
        d.synchronize()
        el = time.perf_counter() - t0
        if step:
            log(f"[perf] {step} steps in {el:.3f} s: {NX * NY * NZ * step / max(el, 1e-12):.4e} zone-updates/s")
        pr = d.get_prim()
    finally:
        d.close()
    log("Done.")
    return pr, t, step


def main(argv=None):
    s = load_settings(argv)
    ic.apply_derived_ic_keys(s)

    def log(msg):
        print(msg, flush=True)

    run(s, log)
    return 0


if __name__ == "__main__":
    sys.exit(main())
