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
order, 1505 values).  Training fits r^ = (y - mur) / s with the mean squared
error, in float32 (GPU) or float64 (here).

`predict` restates the deployed arithmetic of the engine's matrix-core
kernel exactly as oracle/xf_oracle.c lb_predict does: fp16 scaled inputs and
layer-1 activations, 16-term matrix products summed as two exact groups of 8
with one fp32 rounding each, the layer-2 hidden-unit order of the products'
k-slots, and the output unit's two half sums (see lb_predict for the full
statement).  `loss_and_grad` works in float64 on the smooth model (no fp16
rounding) and is the reference the GPU gradients are checked against
(pinned itself by central finite differences in the tests).
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


def _f16(v, saturate: bool = False) -> np.ndarray:
    """fp32 values rounded to fp16 (nearest even), returned as float32."""
    v = np.asarray(v, np.float32)
    if saturate:
        v = np.minimum(v, np.float32(65504.0))
    return v.astype(np.float16).astype(np.float32)


def _fix(v) -> np.ndarray:
    """Multiples of 2^-48 as exact Python integers (object array)."""
    return np.array([int(a) for a in np.ldexp(np.asarray(v, np.float64), 48)], dtype=object)


def _round_fix(n) -> np.ndarray:
    """Integers on the 2^-48 grid rounded to float32 (nearest even)."""
    out = np.empty(len(n), np.float32)
    for r, a in enumerate(n):
        if a == 0:
            out[r] = 0.0
            continue
        m = abs(a)
        sh = max(m.bit_length() - 24, 0)
        if sh:
            q, rem = m >> sh, m & ((1 << sh) - 1)
            half = 1 << (sh - 1)
            m = q + (1 if rem > half or (rem == half and q & 1) else 0)
        out[r] = np.float32(np.ldexp(float(m), sh - 48)) * (1 if a > 0 else -1)
    return out


def _mfma16(c, a_cols, b_cols) -> np.ndarray:
    """c + sum over 16 k-slots of a * b, as two exact groups of 8 with one
    float32 rounding each; a_cols[k] is a scalar weight, b_cols[k] a column
    of fp16 values (or None for a zero slot)."""
    acc = np.asarray(c, np.float32)
    for g in range(2):
        s = _fix(acc)
        for j in range(8):
            a, b = a_cols[8 * g + j], b_cols[8 * g + j]
            if b is None or a == 0:
                continue
            s = s + int(np.ldexp(np.float64(a), 24)) * _fix(np.ldexp(np.asarray(b, np.float64), -24))
        acc = _round_fix(s)
    return acc


def predict(w, x, k: float) -> np.ndarray:
    """The deployed predictor, rounding exactly like oracle/xf_oracle.c."""
    x = np.asarray(x, np.float32)
    W1, b1, W2, b2, w3, b3 = unpack(as_half(w))
    e = scale_exponent(x)
    xs = _f16(np.ldexp(x, -e[:, None]).astype(np.float32))
    n = x.shape[0]
    h1 = np.empty((n, N_HID), np.float32)
    for o in range(N_HID):
        z = _mfma16(np.full(n, b1[o], np.float32), list(W1[o]) + [0.0] * 4,
                    [xs[:, i] for i in range(N_IN)] + [None] * 4)
        h1[:, o] = _f16(np.maximum(z, np.float32(0)), saturate=True)
    h2 = np.empty((n, N_HID), np.float32)
    for o in range(N_HID):
        acc = np.full(n, b2[o], np.float32)
        for p in range(2):
            units = [16 * p + 8 * (j // 4) + 4 * g + j % 4 for g in range(2) for j in range(8)]
            acc = _mfma16(acc, [W2[o, u] for u in units], [h1[:, u] for u in units])
        h2[:, o] = np.maximum(acc, np.float32(0))
    halves = []
    for g in range(2):
        s = np.zeros(n, np.float32)
        for i in range(16):
            u = 8 * (i // 4) + 4 * g + i % 4
            s = (s + (w3[u] * h2[:, u]).astype(np.float32)).astype(np.float32)
        halves.append(s)
    r = ((halves[0] + halves[1]).astype(np.float32) + b3).astype(np.float32)
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
