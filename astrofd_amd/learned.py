"""Learned con2prim guess: sampling, GPU training and evaluation (no counterpart in the reference).

The network (5 -> 64 -> 64 -> 2, SiLU, BF16 MFMA kernels with FP32 accumulation in
libafd) maps the scaled conserved state of a cell to (log10(p/D)/4, log10 W).  The stage
update turns that into the Newton start Z0 = D W + Gamma/(Gamma-1) p W^2 and runs the
reference's FD-Newton iteration (core/rmhd_core.py:273-290) from it; the acceptance test passes the
result only when |f(Z)| < 1e-12 max(1, tau) within 12 evaluations and the recovery
(core/rmhd_core.py:320-341) is valid, otherwise the plain con2prim runs.

Training pairs come from the solver's own con2prim: primitives are sampled, converted with
the library's p2c and recovered with its FP64 con2prim on the GPU; the network learns the
recovered state, which is where Newton converges (not the sampled primitives: the
reference's p2c and con2prim disagree when v.B != 0).

    python -m astrofd_amd.learned --out astrofd_amd/models/c2p_mlp.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from .domain import Domain

NPARAM = 4674
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "c2p_mlp.npz")
GAMMA = 4.0 / 3.0
P_MAX, V_MAX = 1e4, 0.9999


def init_weights(seed=7):
    """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases in the packed layout
    W1[64][5], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((5, 64), (64, 64), (64, 2)):
        bound = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-bound, bound, (fan_out, fan_in)).ravel())
        parts.append(rng.uniform(-bound, bound, fan_out))
    return np.concatenate(parts).astype(np.float32)


def _unit_vectors(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def sample_primitives(n, seed, gamma=GAMMA, slow_frac=0.5):
    """Primitive RMHD states (n x 9): rho log-uniform [1e-4, 1e2], W uniform [1, 50],
    p/rho log-uniform [1e-4, 1e2], magnetisation sigma = b^2/(rho h) log-uniform [1e-3, 10],
    isotropic directions of v and B, psi = 0.  A fraction `slow_frac` of the states instead
    draws W - 1 log-uniform in [1e-11, 49] (|v| from ~4e-6 up): a uniform W leaves the nearly
    static gas that fills most of a jet's box (|v| < 0.02 has W - 1 < 2e-4) out of the data."""
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-4.0, 2.0, n)
    W = rng.uniform(1.0, 50.0, n)
    slow = rng.uniform(0.0, 1.0, n) < slow_frac
    W[slow] = 1.0 + 10.0 ** rng.uniform(-11.0, np.log10(49.0), int(slow.sum()))
    p = rho * 10.0 ** rng.uniform(-4.0, 2.0, n)
    sigma = 10.0 ** rng.uniform(-3.0, 1.0, n)
    vhat = _unit_vectors(rng, n)
    bhat = _unit_vectors(rng, n)
    vmag = np.sqrt(1.0 - 1.0 / (W * W))
    h = 1.0 + gamma / (gamma - 1.0) * p / rho
    cos = np.sum(vhat * bhat, axis=1)
    # b^2 = B^2 / W^2 + (v.B)^2 = |B|^2 (1/W^2 + v^2 cos^2)
    Bmag = np.sqrt(sigma * rho * h / (1.0 / (W * W) + vmag * vmag * cos * cos))
    q = np.zeros((n, 9))
    q[:, 0] = rho
    q[:, 1:4] = vhat * vmag[:, None]
    q[:, 4] = p
    q[:, 5:8] = bhat * Bmag[:, None]
    return q


def make_context(device=0, learned=False):
    """A minimal RMHD context (4^3 cells) carrying the pointwise kernels and the network."""
    p = L.Params()
    p.physics, p.nv, p.nx, p.ny, p.nz, p.ng = L.AFD_PHYS_RMHD, 9, 4, 4, 4, 3
    p.dx = p.dy = p.dz = 0.25
    p.gamma, p.p_max, p.v_max = GAMMA, P_MAX, V_MAX
    p.glm_ch, p.glm_cp = 1.0, 0.1
    p.limiter_id, p.recon_id, p.riemann_id, p.rk_order = 0, 2, 2, 3
    p.precision = L.AFD_PREC_FP64
    p.c2p_guess = L.AFD_C2P_LEARNED if learned else L.AFD_C2P_PLAIN
    p.bc_y = p.bc_z = L.AFD_BC_PERIODIC
    p.device = device
    return Domain(p)


def build_dataset(dom, n, seed, chunk=1 << 22, slow_frac=0.5):
    """(X, Y) of valid pairs from n sampled primitives, and the valid fraction."""
    Xs, Ys, nvalid = [], [], 0
    for off in range(0, n, chunk):
        m = min(chunk, n - off)
        q = sample_primitives(m, seed + off // chunk, slow_frac=slow_frac)
        X, Y, ok = dom.mlp_dataset_from_prims(q)
        Xs.append(X[ok])
        Ys.append(Y[ok])
        nvalid += int(ok.sum())
    return np.concatenate(Xs), np.concatenate(Ys), nvalid / n


def train(dom, X, Y, epochs=200, batch=32768, lr0=3e-3, lr_min=3e-5, seed=11, w0=None, log=print):
    """Adam on the GPU with a cosine learning-rate schedule (one lr per epoch).
    Returns (weights, per-step loss history)."""
    rng = np.random.default_rng(seed)
    dom.set_mlp_weights(init_weights() if w0 is None else w0)
    hist = []
    nbat = X.shape[0] // batch
    for ep in range(epochs):
        perm = rng.permutation(X.shape[0])
        dom.mlp_set_dataset(X[perm], Y[perm])
        lr = lr_min + 0.5 * (lr0 - lr_min) * (1.0 + np.cos(np.pi * ep / max(epochs - 1, 1)))
        h = dom.mlp_train(nbat, batch, lr)
        hist.append(h)
        log(f"epoch {ep + 1}/{epochs} lr {lr:.2e} loss {h.mean():.3e} (last {h[-1]:.3e})")
    return dom.get_mlp_weights(), np.concatenate(hist)


def evaluate(w, n=1 << 20, seed=99, device=0, return_states=False):
    """A/B of plain and learned con2prim on held-out sampled states (the same states for both).

    `newton_valid` marks the states where the plain con2prim converges by Newton to a valid
    state; elsewhere it returns the bisection result or the hydro guess, which no initial guess
    changes, so the learned path can only gain there.  Iterations are the solver's own counts:
    hydro fixed-point iterations + Newton f(Z) evaluations (plain), Newton evaluations from the
    network's guess (+ the plain solve when the acceptance test fails)."""
    q = sample_primitives(n, seed, slow_frac=0.0)  # the specified distribution only (W uniform in [1, 50])
    with make_context(device) as plain, make_context(device, learned=True) as lrn:
        lrn.set_mlp_weights(w)
        _, _, newton_valid = plain.mlp_dataset_from_prims(q)
        U = plain.p2c(q)
        res = {}
        for name, d in (("plain", plain), ("learned", lrn)):
            d.set_instrumentation(stats=True)
            d.reset_stats()
            out, it = d.c2p(U, with_iters=True)
            res[name] = dict(out=out, iters=it.astype(np.float64), stats=d.stats())
    pl, ln = res["plain"], res["learned"]
    m = newton_valid & np.all(np.isfinite(pl["out"]), axis=1) & np.all(np.isfinite(ln["out"]), axis=1)
    rel_rho = np.abs(ln["out"][m, 0] - pl["out"][m, 0]) / np.abs(pl["out"][m, 0])
    rel_p = np.abs(ln["out"][m, 4] - pl["out"][m, 4]) / np.abs(pl["out"][m, 4])
    same = np.all(ln["out"][~newton_valid] == pl["out"][~newton_valid], axis=1)
    sl, sp = ln["stats"], pl["stats"]
    nv = int(newton_valid.sum())
    summary = {
        "states": n,
        "newton_valid_fraction": nv / n,
        "plain_iters_per_state": float(pl["iters"].mean()),
        "learned_iters_per_state": float(ln["iters"].mean()),
        "plain_iters_newton_valid": float(pl["iters"][newton_valid].mean()) if nv else 0.0,
        "learned_iters_newton_valid": float(ln["iters"][newton_valid].mean()) if nv else 0.0,
        "plain_hydro_iters_per_state": sp["hydro_iters"] / n,
        "plain_newton_evals_per_state": sp["newton_evals"] / n,
        "learned_hydro_iters_per_state": sl["hydro_iters"] / n,
        "learned_newton_evals_per_state": sl["newton_evals"] / n,
        "accept_rate": sl["learned_ok"] / n,
        "accept_rate_newton_valid": sl["learned_ok"] / max(nv, 1),
        "plain_bisections": sp["bisections"], "learned_bisections": sl["bisections"],
        "plain_invalid": sp["invalid"], "learned_invalid": sl["invalid"],
        "identical_outside_newton_valid": float(same.mean()) if same.size else 1.0,
        "rho_rel_diff_max": float(rel_rho.max()) if rel_rho.size else 0.0,
        "rho_rel_diff_gt_1e-8_fraction": float(np.mean(rel_rho > 1e-8)) if rel_rho.size else 0.0,
        "rho_rel_diff_median": float(np.median(rel_rho)) if rel_rho.size else 0.0,
        "p_rel_diff_max": float(rel_p.max()) if rel_p.size else 0.0,
    }
    if return_states:
        return summary, dict(cons=U, plain=pl["out"], learned=ln["out"], newton_valid=newton_valid)
    return summary


def save(path, w, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, w=np.asarray(w, dtype=np.float32), meta=json.dumps(meta or {}))


def load(path=MODEL_PATH):
    """Trained weights (float32, 4674) from an npz written by save()."""
    with np.load(path, allow_pickle=False) as f:
        w = np.asarray(f["w"], dtype=np.float32)
    if w.size != NPARAM:
        raise ValueError(f"{path}: {w.size} parameters, expected {NPARAM}")
    return w


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--samples", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--eval-states", type=int, default=1 << 20)
    ap.add_argument("--slow-frac", type=float, default=0.5, help="fraction of states with log-uniform W - 1")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    with make_context() as dom:
        X, Y, frac = build_dataset(dom, a.samples, a.seed, slow_frac=a.slow_frac)
        log(f"dataset: {X.shape[0]} valid pairs of {a.samples} ({frac:.4%}) in {time.perf_counter() - t0:.1f} s")
        t1 = time.perf_counter()
        w, hist = train(dom, X, Y, epochs=a.epochs, batch=a.batch, lr0=a.lr, log=log)
        t_train = time.perf_counter() - t1
    log(f"training: {hist.size} Adam steps in {t_train:.1f} s")
    ev = evaluate(w, n=a.eval_states)
    log(json.dumps(ev))
    meta = dict(samples=a.samples, slow_frac=a.slow_frac, valid_fraction=frac, epochs=a.epochs, batch=a.batch, lr=a.lr,
                seed=a.seed, steps=int(hist.size), final_loss=float(hist[-256:].mean()),
                train_seconds=t_train, eval=ev)
    save(a.out, w, meta)
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
