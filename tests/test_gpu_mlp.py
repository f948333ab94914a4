"""GPU tests of the learned con2prim guess: the BF16-MFMA forward and backward kernels against
the CPU oracle (oracle/mlp.py, pinned by torch autograd and finite differences), the GPU Adam
step, the training loop, the dataset kernel against the C oracle's con2prim, and the A/B of
plain against learned con2prim (what the trained network gains, and how often the acceptance test
rejects it)."""
import os

import numpy as np
import pytest

import oracle
from astrofd_amd import learned
from oracle import mlp
from tests.gpu_helpers import log, make_domain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    d = learned.make_context()
    yield d
    d.close()


def _inputs(n, seed):
    """Feature vectors spread like the scaled states: |x| <= 1.2."""
    rng = np.random.default_rng(seed)
    X = np.empty((n, 5), dtype=np.float32)
    X[:, 0] = rng.uniform(-1.0, 0.5, n)
    X[:, 1] = rng.uniform(-1.2, 0.8, n)
    X[:, 2] = rng.uniform(-2.5, 0.6, n)
    X[:, 3] = rng.uniform(-2.5, 0.8, n)
    X[:, 4] = rng.uniform(-1.0, 1.0, n)
    Y = np.stack([rng.uniform(-1.0, 0.5, n), rng.uniform(0.0, 1.7, n)], axis=1).astype(np.float32)
    return X, Y


def _scaled_weights(seed, gain=2.0):
    """Initial weights scaled up so that every layer carries O(1) signals."""
    return (mlp.init_weights(seed) * gain).astype(np.float32)


@pytest.mark.parametrize("n", [1, 63, 4099])
def test_forward_vs_oracle(ctx, n):
    log(f"test_forward_vs_oracle n={n}")
    w = _scaled_weights(3)
    ctx.set_mlp_weights(w)
    X, _ = _inputs(n, 10 + n)
    y = ctx.mlp_forward(X)
    ref, _ = mlp.forward(w.astype(np.float64), X.astype(np.float64), emulate=True)
    ref64, _ = mlp.forward(w.astype(np.float64), X.astype(np.float64), emulate=False)
    d = np.abs(y - ref)
    d64 = np.abs(y - ref64).max()
    far = np.mean(d > 1e-5)
    log(f"  max |gpu - oracle(bf16)| {d.max():.3e}  mean {d.mean():.3e}  fraction > 1e-5: {far:.2e}  "
        f"|gpu - fp64 net| {d64:.3e}")
    # Outputs agree to FP32 accumulation error (~1e-7) except where the accumulation order or
    # the hardware exp/rcp moves the BF16 rounding of a hidden activation by one BF16 ulp
    # (2^-8 relative); through the next layers that moves an output by a few 1e-3 at most,
    # in a small fraction of samples.  A wrong operand layout or index breaks nearly all.
    assert d.max() <= 5e-3
    assert d.mean() <= 2e-5
    assert far <= 0.02


@pytest.mark.parametrize("n", [64, 8229])
def test_gradient_vs_oracle(ctx, n):
    log(f"test_gradient_vs_oracle n={n}")
    w = _scaled_weights(4)
    ctx.set_mlp_weights(w)
    X, Y = _inputs(n, 20 + n)
    loss, g = ctx.mlp_loss_grad(X, Y)
    lref, gref = mlp.loss_grad(w.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), emulate=True)
    _, g64 = mlp.loss_grad(w.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), emulate=False)
    assert abs(loss - lref) <= 1e-5 * lref, (loss, lref)
    parts = [("W1", mlp.OFF_W1, mlp.OFF_B1), ("b1", mlp.OFF_B1, mlp.OFF_W2), ("W2", mlp.OFF_W2, mlp.OFF_B2),
             ("b2", mlp.OFF_B2, mlp.OFF_W3), ("W3", mlp.OFF_W3, mlp.OFF_B3), ("b3", mlp.OFF_B3, mlp.NPARAM)]
    for name, a, b in parts:
        r = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        r64 = np.linalg.norm(g[a:b] - g64[a:b]) / np.linalg.norm(g64[a:b])
        log(f"  {name}: |g - oracle(bf16)| / |oracle| = {r:.2e}   vs fp64 net {r64:.2e}")
        assert r <= 2e-3, name


def test_adam_step_on_gpu(ctx):
    log("test_adam_step_on_gpu")
    w = _scaled_weights(5)
    X, Y = _inputs(4096, 30)
    ctx.set_mlp_weights(w)
    _, g = ctx.mlp_loss_grad(X, Y)
    ctx.set_mlp_weights(w)  # resets Adam
    ctx.mlp_set_dataset(X, Y)
    hist = ctx.mlp_train(1, 4096, 1e-3)
    w1 = ctx.get_mlp_weights()
    wr, _, _ = mlp.adam(w.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1, 1e-3)
    log(f"  loss {hist[0]:.4e}, max |w_gpu - adam(oracle)| {np.abs(w1 - wr).max():.2e}")
    # the first bias-corrected Adam step is lr * g / (|g| + eps): FP32 rounding only
    np.testing.assert_allclose(w1, wr, rtol=0, atol=2e-7)


def test_training_reduces_loss(ctx):
    log("test_training_reduces_loss")
    q = learned.sample_primitives(1 << 16, 41)
    X, Y, ok = ctx.mlp_dataset_from_prims(q)
    X, Y = X[ok], Y[ok]
    ctx.set_mlp_weights(learned.init_weights())
    ctx.mlp_set_dataset(X, Y)
    hist = ctx.mlp_train(400, 4096, 3e-3)
    log(f"  valid {ok.mean():.4f}; loss first {hist[:5].mean():.3e} last {hist[-5:].mean():.3e}")
    assert np.all(np.isfinite(hist))
    assert hist[-10:].mean() < 0.3 * hist[:5].mean()


def test_dataset_vs_oracle(ctx):
    log("test_dataset_vs_oracle")
    q = learned.sample_primitives(3000, 42)
    X, Y, ok = ctx.mlp_dataset_from_prims(q)
    P = oracle.make_params("rmhd", gamma=learned.GAMMA, p_max=learned.P_MAX, v_max=learned.V_MAX)
    U = oracle.rmhd_p2c(P, q)
    qo, flags = oracle.rmhd_c2p_flags(P, U)
    Xo, oko = mlp.features(U)
    Yo = mlp.targets(qo, U)
    # pairs only where the con2prim converged by Newton to a valid state
    oko &= np.all(np.isfinite(Yo), axis=1) & ((flags & 6) == 0)
    log(f"  valid gpu {ok.mean():.4f} oracle {oko.mean():.4f}")
    assert np.array_equal(ok, oko)
    # device features use the FP32 log2 instruction; targets come from the FP64 con2prim
    np.testing.assert_allclose(X[ok], Xo[ok], rtol=0, atol=2e-6)
    np.testing.assert_allclose(Y[ok, 1], Yo[ok, 1], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(Y[ok, 0], Yo[ok, 0], rtol=0, atol=1e-6)


def _trained_weights():
    if os.path.exists(learned.MODEL_PATH):
        return learned.load()
    pytest.skip(f"{learned.MODEL_PATH} is missing: train it with `python -m astrofd_amd.learned`")


def _residual(P, U, q, gamma):
    """|f(Z)| / max(1, tau) of recovered primitives q for conserved states U (the reference's
    convergence measure, core/rmhd_core.py:273-290), with Z = rho h W^2 from q."""
    v2 = np.sum(q[:, 1:4] ** 2, axis=1)
    Z = q[:, 0] * (1.0 + gamma / (gamma - 1.0) * q[:, 4] / q[:, 0]) / (1.0 - v2)
    return np.abs(oracle.rmhd_f_of_Z(P, Z, U)) / np.maximum(1.0, U[:, 4])


def test_learned_guess_ab():
    """What the trained network gains: the same held-out states through plain and learned
    con2prim.  Where the plain solve converges by Newton to a valid state, the acceptance test must pass
    the learned start for nearly every state and the learned path must need fewer iterations;
    elsewhere (bisection, hydro-guess fallback) both paths return the same state."""
    log("test_learned_guess_ab")
    ev, st = learned.evaluate(_trained_weights(), n=1 << 18, seed=123, return_states=True)
    for k, v in ev.items():
        log(f"  {k}: {v}")
    assert ev["accept_rate_newton_valid"] >= 0.95
    assert ev["learned_iters_newton_valid"] < 0.6 * ev["plain_iters_newton_valid"]
    assert ev["learned_iters_per_state"] < ev["plain_iters_per_state"]
    assert ev["identical_outside_newton_valid"] >= 0.99
    assert ev["rho_rel_diff_median"] <= 1e-10  # the same root
    # The plain stop test |f| < 1e-10 max(1, tau) fixes Z only to that over |f'(Z)|, loose where f
    # is flat, so single states may differ more; the learned result must be at least as converged.
    m = st["newton_valid"]
    P = oracle.make_params("rmhd", gamma=learned.GAMMA, p_max=learned.P_MAX, v_max=learned.V_MAX)
    U = st["cons"][m]
    r_pl = _residual(P, U, st["plain"][m], learned.GAMMA)
    r_ln = _residual(P, U, st["learned"][m], learned.GAMMA)
    log(f"  residual |f|/max(1,tau) median plain {np.median(r_pl):.2e} learned {np.median(r_ln):.2e}; "
        f"99th pct plain {np.percentile(r_pl, 99):.2e} learned {np.percentile(r_ln, 99):.2e}")
    assert np.median(r_ln) <= np.median(r_pl)
    assert np.percentile(r_ln, 99) <= max(np.percentile(r_pl, 99), 1e-10)


def test_learned_step_matches_plain():
    """A full RMHD step (WENO5 + HLLD + SSPRK3) with the learned guess agrees with the plain one."""
    log("test_learned_step_matches_plain")
    from tests.golden_states import smooth_state
    w = _trained_weights()
    ng = 3
    pr = smooth_state((20 + 2 * ng, 18 + 2 * ng, 16 + 2 * ng), 77, "rmhd", vamp=0.5)
    outs = []
    for guess in ("plain", "learned"):
        d = make_domain("rmhd", pr.shape, ng, gamma=4.0 / 3.0, recon="weno", riemann="hlld_full", rk_order=3,
                        c2p_guess=guess)
        if guess == "learned":
            d.set_mlp_weights(w)
        d.set_instrumentation(stats=True)
        d.set_prim(pr)
        d.step(1e-3)
        outs.append(d.get_prim())
        st = d.stats()
        log(f"  {guess}: {st}")
        d.close()
    core = (slice(None),) + (slice(ng, -ng),) * 3
    a, b = outs[0][core], outs[1][core]
    r = np.abs(a - b).max() / np.abs(a).max()
    log(f"  max rel diff {r:.3e}")
    assert r <= 1e-9


def test_previous_state_guess_matches_plain():
    """The previous-state Newton start (C2P_GUESS = previous) agrees with the plain con2prim."""
    log("test_previous_state_guess_matches_plain")
    from tests.golden_states import smooth_state
    ng = 3
    pr = smooth_state((20 + 2 * ng, 18 + 2 * ng, 16 + 2 * ng), 78, "rmhd", vamp=0.5)
    outs = []
    for guess in ("plain", "previous"):
        d = make_domain("rmhd", pr.shape, ng, gamma=4.0 / 3.0, recon="weno", riemann="hlld_full", rk_order=3,
                        c2p_guess=guess)
        d.set_instrumentation(stats=True)
        d.set_prim(pr)
        d.step(1e-3)
        outs.append(d.get_prim())
        log(f"  {guess}: {d.stats()}")
        d.close()
    core = (slice(None),) + (slice(ng, -ng),) * 3
    r = np.abs(outs[0][core] - outs[1][core]).max() / np.abs(outs[0][core]).max()
    log(f"  max rel diff {r:.3e}")
    assert r <= 1e-9


@pytest.mark.parametrize("kind", ["zero", "nan"])
def test_bad_network_falls_back_to_plain(kind):
    """A useless network (all weights zero) or a broken one (all NaN) cannot change the result: the
    acceptance test rejects every start it cannot confirm, those cells take the plain con2prim,
    and the step equals the plain step.  With NaN weights no start is accepted."""
    log(f"test_bad_network_falls_back_to_plain {kind}")
    from tests.golden_states import smooth_state
    ng = 3
    pr = smooth_state((20 + 2 * ng, 18 + 2 * ng, 16 + 2 * ng), 79, "rmhd", vamp=0.5)
    w = np.zeros(learned.NPARAM, dtype=np.float32)
    if kind == "nan":
        w[:] = np.nan
    outs, stats = [], []
    for guess in ("plain", "learned"):
        d = make_domain("rmhd", pr.shape, ng, gamma=4.0 / 3.0, recon="weno", riemann="hlld_full", rk_order=3,
                        c2p_guess=guess)
        if guess == "learned":
            d.set_mlp_weights(w)
        d.set_instrumentation(stats=True)
        d.set_prim(pr)
        d.step(1e-3)
        outs.append(d.get_prim())
        stats.append(d.stats())
        log(f"  {guess}: {stats[-1]}")
        d.close()
    core = (slice(None),) + (slice(ng, -ng),) * 3
    a, b = outs[0][core], outs[1][core]
    assert np.all(np.isfinite(b))
    r = np.abs(a - b).max() / np.abs(a).max()
    log(f"  max rel diff {r:.3e}")
    assert r <= 1e-9
    if kind == "nan":
        assert stats[1]["learned_ok"] == 0
