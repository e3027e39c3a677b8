"""numpy restatement of the learned absorbing layer's column model and its
training loss -- TEST INFRASTRUCTURE ONLY (the checker of the HIP trainer and
of the engine's column predictor; the product never imports it).

The model (xf_learned_boundary in include/xfdtd_hip.h; no reference
counterpart): for a column x of 12 inputs (u^{t+1} on layers 1..3, u^t on
layers 0..3, u^{t-1} on layers 0..2, the in-plane Laplacians of u^t on
layers 1 and 0)

    s    = 2^e, the power of two with max|x_i| in [2^(e-1), 2^e)
    mur  = x4 + k (x0 - x3)
    pred = mur + s * MLP(x / s)
    MLP  = w3 . relu(W2 relu(W1 x^ + b1) + b2) + b3

with fp16 weights (W1 32x12, b1 32, W2 32x32, b2 32, w3 32, b3 1, in that
order, 1505 values) and float32 arithmetic, one rounding per product and per
sum, sums in input order.  Training fits r^ = (y - mur) / s with the mean
squared error.

`predict` reproduces the float32 roundings of the C/HIP predictor exactly;
`loss_and_grad` works in float64 and is the reference the GPU gradients are
checked against (pinned itself by central finite differences in the tests).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 12, 32
N_PARAM = N_HID * N_IN + N_HID + N_HID * N_HID + N_HID + N_HID + 1


def unpack(w):
    """Split a flat parameter vector into (W1, b1, W2, b2, w3, b3)."""
    w = np.asarray(w)
    o = 0
    W1 = w[o:o + N_HID * N_IN].reshape(N_HID, N_IN); o += N_HID * N_IN
    b1 = w[o:o + N_HID]; o += N_HID
    W2 = w[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = w[o:o + N_HID]; o += N_HID
    w3 = w[o:o + N_HID]; o += N_HID
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def as_half(w) -> np.ndarray:
    """Weights rounded to fp16, returned as float32 values."""
    w = np.asarray(w)
    if w.dtype == np.uint16:
        return w.view(np.float16).astype(np.float32)
    return np.asarray(w, np.float32).astype(np.float16).astype(np.float32)


def scale_exponent(x: np.ndarray) -> np.ndarray:
    """e with max|x| in [2^(e-1), 2^e) per row (0 where the row is zero)."""
    m = np.max(np.abs(x), axis=1)
    _, e = np.frexp(m)
    return np.where(m > 0, e, 0).astype(np.int32)


def mur(x: np.ndarray, k: float) -> np.ndarray:
    x = np.asarray(x, np.float32)
    k = np.float32(k)
    return (x[:, 4] + k * (x[:, 0] - x[:, 3])).astype(np.float32)


def predict(w, x, k: float) -> np.ndarray:
    """The float32 predictor, rounding exactly like the C/HIP code."""
    x = np.asarray(x, np.float32)
    W1, b1, W2, b2, w3, b3 = unpack(as_half(w))
    e = scale_exponent(x)
    xs = np.ldexp(x, -e[:, None]).astype(np.float32)
    n = x.shape[0]
    h1 = np.empty((n, N_HID), np.float32)
    for o in range(N_HID):
        acc = np.full(n, b1[o], np.float32)
        for i in range(N_IN):
            acc = (acc + W1[o, i] * xs[:, i]).astype(np.float32)
        h1[:, o] = np.maximum(acc, np.float32(0))
    r = np.full(n, b3, np.float32)
    for o in range(N_HID):
        acc = np.full(n, b2[o], np.float32)
        for i in range(N_HID):
            acc = (acc + W2[o, i] * h1[:, i]).astype(np.float32)
        r = (r + w3[o] * np.maximum(acc, np.float32(0))).astype(np.float32)
    m = mur(x, k)
    out = (m + np.ldexp(r, e).astype(np.float32)).astype(np.float32)
    return np.where(np.max(np.abs(x), axis=1) > 0, out, m)


def targets(samples: np.ndarray, k: float, drop_zero: bool = True):
    """(x^, r^) training pairs from raw samples (n, 13): inputs scaled by s,
    residual target (y - mur) / s.  Rows with s = 0 are dropped, or kept as
    (0, 0) pairs (the HIP trainer's convention) with drop_zero=False."""
    x = np.asarray(samples[:, :N_IN], np.float32)
    y = np.asarray(samples[:, N_IN], np.float32)
    e = scale_exponent(x)
    keep = np.max(np.abs(x), axis=1) > 0
    xs = np.ldexp(x, -e[:, None]).astype(np.float32)
    r = np.ldexp((y - mur(x, k)).astype(np.float32), -e).astype(np.float32)
    if not drop_zero:
        return np.where(keep[:, None], xs, 0).astype(np.float32), np.where(keep, r, 0).astype(np.float32)
    return xs[keep], r[keep]


def loss_and_grad(w, xs, r):
    """Mean squared error of the MLP on scaled inputs xs against r, and its
    gradient with respect to the (fp16-rounded) weights, in float64."""
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, np.float64) for a in unpack(as_half(w)))
    xs = np.asarray(xs, np.float64)
    r = np.asarray(r, np.float64)
    n = xs.shape[0]
    z1 = xs @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ w3 + b3
    d = out - r
    loss = float(np.mean(d * d))
    g_out = 2.0 * d / n
    g_w3 = h2.T @ g_out
    g_b3 = g_out.sum()
    g_z2 = np.outer(g_out, w3) * (z2 > 0)
    g_W2 = g_z2.T @ h1
    g_b2 = g_z2.sum(0)
    g_z1 = (g_z2 @ W2) * (z1 > 0)
    g_W1 = g_z1.T @ xs
    g_b1 = g_z1.sum(0)
    grad = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, grad


def loss_only(w64, xs, r):
    """MSE with float64 weights taken as given (no fp16 rounding): the
    function whose derivative loss_and_grad returns, for finite differences."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(w64, np.float64))
    h1 = np.maximum(np.asarray(xs, np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    d = h2 @ w3 + b3 - np.asarray(r, np.float64)
    return float(np.mean(d * d))


def init_weights(seed: int = 0, out_scale: float = 0.0) -> np.ndarray:
    """He-initialised hidden layers; the output layer starts at out_scale x
    that (0: the model starts as the plain Mur update)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / N_IN), (N_HID, N_IN))
    W2 = rng.normal(0, np.sqrt(2 / N_HID), (N_HID, N_HID))
    w3 = rng.normal(0, np.sqrt(1 / N_HID), N_HID) * out_scale
    return np.concatenate([W1.ravel(), np.zeros(N_HID), W2.ravel(), np.zeros(N_HID), w3,
                           [0.0]]).astype(np.float32)
