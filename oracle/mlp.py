"""CPU oracle of the learned con2prim guess network.

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU leg of the
benchmark use this module as the checker of the GPU kernels; nothing in the
product imports it.

The network (no counterpart in the reference): 5 -> 64 -> 64 -> 2 with SiLU,
parameters in one FP32 vector  W1[64][5], b1[64], W2[64][64], b2[64],
W3[2][64], b3[2]  (4674 values).  Loss = mean over samples of the summed
squared error of the two outputs.

`emulate=True` reproduces where the GPU kernels round operands to BF16
(weights, inputs, hidden activations, back-propagated deltas; accumulation is
FP32, here FP64; the biases b1 and b2 enter as the pair hi + lo of BF16 values,
b1 and b2 of 16 significant bits, and b3 in FP32), so that the GPU result
differs from the oracle only by FP32 accumulation order and the hardware
exp/reciprocal.
`emulate=False` is the plain FP64 network, pinned in the tests by central
finite differences of the loss and by torch autograd on the CPU.

The features and targets restate the device code of the library: the
conserved state (D, S, tau, B) maps to five O(1) numbers
  x0 = log10 D / 4, x1 = log10(tau / D) / 4, x2 = log10(|S| / D + 1e-10) / 4,
  x3 = log10(B^2 / D + 1e-10) / 4, x4 = S.B / (|S| |B|),
and the targets are y0 = log10(p / D) / 4, y1 = log10 W of the con2prim output.
"""
import numpy as np

NIN, NH, NOUT = 5, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT  # 4674


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    assert w.size == NPARAM
    return (w[OFF_W1:OFF_B1].reshape(NH, NIN), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(NH, NH),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3].reshape(NOUT, NH), w[OFF_B3:])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(a) for a in (W1, b1, W2, b2, W3, b3)])


def bf16(a):
    """Round to the nearest BF16 value, ties to even (through FP32), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def _ident(a):
    return np.asarray(a, dtype=np.float64)


def bias_split(b):
    """hi + lo with hi = bf16(b) and lo = bf16(b - hi): the biases as the GPU's MFMAs see them."""
    b = np.asarray(b, dtype=np.float32).astype(np.float64)
    hi = bf16(b)
    return hi + bf16(b - hi)


def sigmoid(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(w, X, emulate=True):
    """Outputs (n x 2) and the intermediates of the forward pass."""
    r = bf16 if emulate else _ident
    rb = bias_split if emulate else _ident
    W1, b1, W2, b2, W3, b3 = unpack(w)
    xb = r(X)                                  # n x 5
    z1 = xb @ r(W1).T + rb(b1)                 # n x 64
    h1 = r(silu(z1))
    z2 = h1 @ r(W2).T + rb(b2)
    h2 = r(silu(z2))
    o = h2 @ r(W3).T + b3
    return o, dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2)


def loss_grad(w, X, Y, emulate=True):
    """(loss, gradient in the packed parameter layout) of the mean squared error."""
    r = bf16 if emulate else _ident
    W1, b1, W2, b2, W3, b3 = unpack(w)
    n = X.shape[0]
    o, t = forward(w, X, emulate)
    e = o - np.asarray(Y, dtype=np.float64)
    loss = float(np.sum(e * e) / n)
    d = 2.0 * e / n                           # dL/do
    db = r(d)
    dz2 = (db @ r(W3)) * dsilu(t["z2"])
    dz2b = r(dz2)
    dz1 = (dz2b @ r(W2)) * dsilu(t["z1"])
    dz1b = r(dz1)
    g = pack(dz1b.T @ t["xb"], dz1.sum(0), dz2b.T @ t["h1"], dz2.sum(0), db.T @ t["h2"], d.sum(0))
    return loss, g


def adam(w, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected) -> (w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    w = w - lr * (m / (1.0 - beta1 ** t)) / (np.sqrt(v / (1.0 - beta2 ** t)) + eps)
    return w, m, v


def init_weights(seed=7):
    """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, FP32."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, NOUT)):
        bound = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-bound, bound, (fan_out, fan_in)))
        parts.append(rng.uniform(-bound, bound, fan_out))
    return pack(*parts).astype(np.float32)


def features(U):
    """Scaled features of conserved states (n x 9, FP64): (X n x 5, ok n)."""
    U = np.asarray(U, dtype=np.float64)
    D, tau = U[:, 0], U[:, 4]
    S2 = np.sum(U[:, 1:4] ** 2, axis=1)
    B2 = np.sum(U[:, 5:8] ** 2, axis=1)
    SB = np.sum(U[:, 1:4] * U[:, 5:8], axis=1)
    ok = (D > 0) & (tau > 0) & np.isfinite(D) & np.isfinite(tau) & np.isfinite(S2) & np.isfinite(B2)
    with np.errstate(all="ignore"):
        X = np.stack([np.log10(D) / 4, np.log10(tau / D) / 4, np.log10(np.sqrt(S2) / D + 1e-10) / 4,
                      np.log10(B2 / D + 1e-10) / 4,
                      np.where(S2 * B2 > 0, SB / np.sqrt(S2 * B2), 0.0)], axis=1)
    ok &= np.all(np.isfinite(X), axis=1)
    X[~ok] = 0.0
    return X, ok


def targets(q, U):
    """(log10(p / D) / 4, log10 W) of primitive states q for conserved states U."""
    v2 = np.minimum(np.sum(q[:, 1:4] ** 2, axis=1), 1.0 - 1e-14)
    W = 1.0 / np.sqrt(1.0 - v2)
    with np.errstate(all="ignore"):
        return np.stack([np.log10(q[:, 4] / U[:, 0]) / 4, np.log10(W)], axis=1)


def guess_Z(D, gfac, y):
    """Z = rho h W^2 = D W + gfac p W^2 from the network outputs."""
    p = D * 10.0 ** (4.0 * y[:, 0])
    W = np.maximum(10.0 ** y[:, 1], 1.0)
    return D * W + gfac * p * W * W
