"""Pins the CPU oracle of the learned-guess network (oracle/mlp.py) before the GPU is checked against it.

The network has no counterpart in the reference, so the checks come from outside the
oracle's own formulas: torch autograd on the CPU (an independent implementation of the
same network), central finite differences of the loss, torch.optim.Adam, torch's BF16
rounding, and closed forms of the sampling and the features.
"""
import numpy as np
import pytest
import torch

from astrofd_amd import learned
from oracle import mlp


def _weights(seed=3):
    return mlp.init_weights(seed).astype(np.float64)


def _data(n, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1.0, 1.0, (n, 5))
    Y = rng.uniform(-0.5, 1.5, (n, 2))
    return X, Y


def _torch_net(w):
    W1, b1, W2, b2, W3, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in mlp.unpack(w))
    return [W1, b1, W2, b2, W3, b3]


def _torch_loss(params, X, Y):
    W1, b1, W2, b2, W3, b3 = params
    x = torch.tensor(X, dtype=torch.float64)
    h1 = torch.nn.functional.silu(x @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    o = h2 @ W3.T + b3
    return ((o - torch.tensor(Y)) ** 2).sum() / X.shape[0], o


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-30, 30, 20000),
                        # exact ties between two BF16 values: both directions of the even rule
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0])])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)
    assert mlp.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0            # tie -> even (down)
    assert mlp.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6  # tie -> even (up)


def test_forward_matches_torch():
    w = _weights()
    X, Y = _data(257)
    o, _ = mlp.forward(w, X, emulate=False)
    _, ot = _torch_loss(_torch_net(w), X, Y)
    np.testing.assert_allclose(o, ot.detach().numpy(), rtol=1e-13, atol=1e-13)


def test_gradient_matches_torch_autograd():
    w = _weights()
    X, Y = _data(300)
    loss, g = mlp.loss_grad(w, X, Y, emulate=False)
    params = _torch_net(w)
    lt, _ = _torch_loss(params, X, Y)
    lt.backward()
    gt = np.concatenate([p.grad.numpy().ravel() for p in params])
    assert abs(loss - lt.item()) <= 1e-13 * abs(loss)
    np.testing.assert_allclose(g, gt, rtol=1e-11, atol=1e-14)


def test_gradient_matches_finite_differences():
    w = _weights()
    X, Y = _data(64)
    _, g = mlp.loss_grad(w, X, Y, emulate=False)
    rng = np.random.default_rng(1)
    # every bias, and weights of every matrix (first/last entries + random ones)
    idx = np.unique(np.concatenate([
        np.arange(mlp.OFF_B1, mlp.OFF_W2), np.arange(mlp.OFF_B2, mlp.OFF_W3), np.arange(mlp.OFF_B3, mlp.NPARAM),
        [mlp.OFF_W1, mlp.OFF_B1 - 1, mlp.OFF_W2, mlp.OFF_B2 - 1, mlp.OFF_W3, mlp.OFF_B3 - 1],
        rng.integers(0, mlp.NPARAM, 200)]))
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss_grad(wp, X, Y, emulate=False)[0] - mlp.loss_grad(wm, X, Y, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_emulation_changes_results_at_bf16_level():
    # the emulated network differs from the FP64 one by the BF16 rounding, not more
    w = _weights()
    X, Y = _data(1000)
    o64, _ = mlp.forward(w, X, emulate=False)
    o16, _ = mlp.forward(w, X, emulate=True)
    d = np.abs(o16 - o64).max()
    assert 1e-5 < d < 5e-2
    _, g64 = mlp.loss_grad(w, X, Y, emulate=False)
    _, g16 = mlp.loss_grad(w, X, Y, emulate=True)
    r = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < r < 5e-2


def test_adam_matches_torch():
    w = _weights()
    X, Y = _data(128)
    params = [torch.tensor(w.copy(), requires_grad=True)]
    opt = torch.optim.Adam(params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    wo = w.copy()
    for t in range(1, 4):
        _, g = mlp.loss_grad(wo, X, Y, emulate=False)
        wo, m, v = mlp.adam(wo, m, v, g, t, 1e-3)
        opt.zero_grad()
        params[0].grad = torch.tensor(mlp.loss_grad(params[0].detach().numpy(), X, Y, emulate=False)[1])
        opt.step()
    np.testing.assert_allclose(wo, params[0].detach().numpy(), rtol=0, atol=1e-12)


def test_init_weights_agree():
    np.testing.assert_array_equal(learned.init_weights(7), mlp.init_weights(7))


def test_sampled_primitives_have_the_stated_distribution():
    g = learned.GAMMA
    q = learned.sample_primitives(20000, 4)
    rho, v, p, B = q[:, 0], q[:, 1:4], q[:, 4], q[:, 5:8]
    v2 = np.sum(v * v, axis=1)
    W = 1.0 / np.sqrt(1.0 - v2)
    h = 1.0 + g / (g - 1.0) * p / rho
    b2 = np.sum(B * B, axis=1) / W ** 2 + np.sum(v * B, axis=1) ** 2
    sigma = b2 / (rho * h)
    assert np.all((W >= 1.0 - 1e-12) & (W <= 50.0 + 1e-9))
    assert np.all((np.log10(rho) >= -4) & (np.log10(rho) <= 2))
    assert np.all((np.log10(p / rho) >= -4 - 1e-12) & (np.log10(p / rho) <= 2 + 1e-12))
    assert np.all((np.log10(sigma) >= -3 - 1e-9) & (np.log10(sigma) <= 1 + 1e-9))
    # log-uniform in sigma: the mean of log10 sigma is the midpoint of [-3, 1]
    assert abs(np.mean(np.log10(sigma)) + 1.0) < 0.05
    assert np.all(q[:, 8] == 0.0)


def test_features_closed_form():
    U = np.zeros((3, 9))
    U[0] = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 0]                 # at rest, unmagnetised
    U[1] = [10.0, 3.0, 4.0, 0, 1000.0, 0, 0, 2.0, 0]       # |S| = 5, S.B = 0
    U[2] = [2.0, 1.0, 0, 0, -1.0, 1.0, 0, 0, 0]            # tau < 0: not a valid input
    X, ok = mlp.features(U)
    np.testing.assert_allclose(X[0], [0.0, 0.0, -2.5, -2.5, 0.0], atol=1e-15)
    np.testing.assert_allclose(X[1], [0.25, 0.5, np.log10(0.5 + 1e-10) / 4, np.log10(0.4 + 1e-10) / 4, 0.0],
                               atol=1e-15)
    assert ok.tolist() == [True, True, False]
    y = np.array([[np.log10(0.1) / 4, np.log10(2.0)]])
    Z = mlp.guess_Z(np.array([3.0]), 4.0, y)
    assert abs(Z[0] - (3.0 * 2.0 + 4.0 * 0.3 * 4.0)) < 1e-12
