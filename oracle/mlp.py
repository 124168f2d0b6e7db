"""CPU reference of the learned search-radius model -- TEST INFRASTRUCTURE ONLY.

The network is a build addition with no counterpart in the reference crate:
27 -> 64 -> 64 -> 2 with ReLU, bf16 weights and activations, fp32 (here
float64) accumulation.  This module restates the device kernels' arithmetic
so that tests can check them:

  - ``features``: log(1 + count) of the 3x3x3 grid cells around a point's
    grid cell (x offset outermost, zero outside the grid), rounded to bf16.
    The counts come from the grid's delimiters (celery.rs:372-414).
  - ``forward`` / ``gradient``: forward pass and the gradient of
    0.5 * sum((y - t)^2) / n, with the same bf16 rounding points as the
    device (h1, h2, dz2, dz1 and the weights).  With ``bf16=False`` no
    rounding happens and the function is the plain float64 network, which
    tests pin by finite differences.

Parameter layout (float32 vector of PARAMS): W1[32][64], W2[64][64],
W3[64][16], b1[64], b2[64], b3[16]; weights are [input][output].
"""
from __future__ import annotations

import numpy as np

IN, FEATURES, H, OUT, OUT_PAD = 32, 27, 64, 2, 16
N_W1, N_W2, N_W3 = IN * H, H * H, H * OUT_PAD
WEIGHTS = N_W1 + N_W2 + N_W3
PARAMS = WEIGHTS + H + H + OUT_PAD


def bf16_bits(x) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even (NaN kept quiet)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_value(bits) -> np.ndarray:
    """bf16 bit patterns -> float64 values."""
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16(x) -> np.ndarray:
    return bf16_value(bf16_bits(np.asarray(x, np.float32)))


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + N_W1].reshape(IN, H); o += N_W1
    W2 = p[o:o + N_W2].reshape(H, H); o += N_W2
    W3 = p[o:o + N_W3].reshape(H, OUT_PAD); o += N_W3
    b1 = p[o:o + H]; o += H
    b2 = p[o:o + H]; o += H
    b3 = p[o:o + OUT_PAD]
    return W1, W2, W3, b1, b2, b3


def pack(W1, W2, W3, b1, b2, b3):
    return np.concatenate([np.ravel(a) for a in (W1, W2, W3, b1, b2, b3)])


def features(bins, delim, cpd):
    """bf16 bit patterns [n][32] from each point's grid cell and the delimiters."""
    bins = np.asarray(bins, np.int64)
    delim = np.asarray(delim, np.int64)
    ix, iy, iz = bins // (cpd * cpd), (bins // cpd) % cpd, bins % cpd
    out = np.zeros((len(bins), IN), np.uint16)
    j = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                bx, by, bz = ix + dx, iy + dy, iz + dz
                ok = (bx >= 0) & (bx < cpd) & (by >= 0) & (by < cpd) & (bz >= 0) & (bz < cpd)
                b = np.where(ok, (bx * cpd + by) * cpd + bz, 0)
                c = np.where(ok, delim[b + 1] - delim[b], 0)
                out[:, j] = bf16_bits(np.log1p(c.astype(np.float64)).astype(np.float32))
                j += 1
    return out


def _weights(p, rounded):
    W1, W2, W3, b1, b2, b3 = unpack(p)
    if rounded:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    return W1, W2, W3[:, :OUT], b1, b2, b3[:OUT]


def forward(p, x, rounded=True):
    """Network outputs [n][2] for bf16 inputs x (bit patterns [n][32])."""
    W1, W2, W3, b1, b2, b3 = _weights(p, rounded)
    X = bf16_value(x) if np.asarray(x).dtype == np.uint16 else np.asarray(x, np.float64)
    r = bf16 if rounded else (lambda a: a)
    h1 = r(np.maximum(X @ W1 + b1, 0.0))
    h2 = r(np.maximum(h1 @ W2 + b2, 0.0))
    return h2 @ W3 + b3


def gradient(p, x, t, rounded=True):
    """(loss, gradient vector) of 0.5 * sum((y - t)^2) / n."""
    W1, W2, W3, b1, b2, b3 = _weights(p, rounded)
    X = bf16_value(x) if np.asarray(x).dtype == np.uint16 else np.asarray(x, np.float64)
    T = np.asarray(t, np.float64)
    n = len(X)
    r = bf16 if rounded else (lambda a: a)
    z1 = X @ W1 + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ W2 + b2
    h2 = r(np.maximum(z2, 0.0))
    y = h2 @ W3 + b3
    diff = y - T
    loss = 0.5 * np.sum(diff * diff) / n
    dy = diff / n
    gW3 = np.zeros((H, OUT_PAD))
    gW3[:, :OUT] = h2.T @ dy
    gb3 = np.zeros(OUT_PAD)
    gb3[:OUT] = dy.sum(0)
    dz2 = r(np.where(z2 > 0, dy @ W3.T, 0.0))
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = r(np.where(z1 > 0, dz2 @ W2.T, 0.0))
    gW1 = X.T @ dz1
    gb1 = dz1.sum(0)
    return loss, pack(gW1, gW2, gW3, gb1, gb2, gb3)


def loss_only(p, x, t, rounded=False):
    y = forward(p, x, rounded)
    d = y - np.asarray(t, np.float64)
    return 0.5 * np.sum(d * d) / len(d)


def init_params(seed):
    """Random parameters in the device layout (zero pads), for tests."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((IN, H)); W1[:FEATURES] = rng.uniform(-1, 1, (FEATURES, H)) * np.sqrt(6 / FEATURES)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6 / H)
    W3 = np.zeros((H, OUT_PAD)); W3[:, :OUT] = rng.uniform(-1, 1, (H, OUT)) * np.sqrt(6 / H)
    b1 = rng.normal(0, 0.1, H)
    b2 = rng.normal(0, 0.1, H)
    b3 = np.zeros(OUT_PAD); b3[:OUT] = rng.normal(0, 0.1, OUT)
    return pack(W1, W2, W3, b1, b2, b3).astype(np.float32)
