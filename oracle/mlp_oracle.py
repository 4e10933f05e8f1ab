"""CPU oracle of the learned root-cost model (test infrastructure only).

Imported by tests/ (and nothing in the product path).  A numpy restatement
of the library's BF16 MLP 8 -> 64 -> 64 -> 2 (ReLU), its loss and the
gradient of that loss with respect to the packed parameter vector:

    params = W1[8][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
    norm   = feature mean[8], std[8], target mean[2], std[2]
    x_n    = (x - mean) * (1/std)                          (fp32)
    h1     = relu(bf(x_n) @ bf(W1) + b1)
    h2     = relu(bf(h1) @ bf(W2) + b2)
    out    = bf(h2) @ bf(W3) + b3
    L      = sum((out - y_n)^2) / (2 rows),  y_n = (y - mean)/std

bf() rounds to bfloat16 (nearest, ties to even) where the GPU kernels feed
a matrix product; with bf16=False every bf() is the identity and the
arithmetic is float64, which is what the finite-difference checks use.
The reference (btsun/kclistpp) has no cost model: this restates the
library's own definition, pinned by finite differences, by torch's
bfloat16 rounding and by closed forms in tests/test_costmodel_oracle.py.
"""

from __future__ import annotations

import numpy as np

F, H, O = 8, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * O
P = OFF_B3 + O
NORM = 2 * F + 2 * O


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    p = np.asarray(params, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(F, H)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(H, H)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(H, O)
    b3 = p[OFF_B3:P]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3])


def identity_norm() -> np.ndarray:
    n = np.zeros(NORM, np.float32)
    n[F:2 * F] = 1.0
    n[2 * F + O:] = 1.0
    return n


def _normalise(x, y, norm, bf: bool):
    norm = np.asarray(norm, np.float32)
    mu, sd = norm[:F], norm[F:2 * F]
    ymu, ysd = norm[2 * F:2 * F + O], norm[2 * F + O:]
    if bf:  # the kernels' fp32 arithmetic
        xn = ((np.asarray(x, np.float32) - mu) * (np.float32(1) / sd)).astype(np.float64)
        yn = None if y is None else ((np.asarray(y, np.float32) - ymu) / ysd).astype(np.float64)
    else:
        xn = (np.asarray(x, np.float64) - mu) / sd.astype(np.float64)
        yn = None if y is None else (np.asarray(y, np.float64) - ymu) / ysd.astype(np.float64)
    return xn, yn


def forward(params, norm, x, bf: bool = True):
    """Normalised outputs (rows, 2) and the activations the backward needs."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params)
    xn, _ = _normalise(x, None, norm, bf)
    xb = r(xn)
    h1 = np.maximum(xb @ r(W1) + b1, 0.0)
    h1b = r(h1)
    h2 = np.maximum(h1b @ r(W2) + b2, 0.0)
    h2b = r(h2)
    out = h2b @ r(W3) + b3
    return out, (xb, h1, h1b, h2, h2b)


def predict(params, norm, x, bf: bool = True) -> np.ndarray:
    """De-normalised predictions (log2(1 + cycles) of build and counting)."""
    out, _ = forward(params, norm, x, bf)
    norm = np.asarray(norm, np.float32).astype(np.float64)
    return out * norm[2 * F + O:] + norm[2 * F:2 * F + O]


def loss_and_grad(params, norm, x, y, bf: bool = True, plant: str | None = None):
    """L = sum((out - y_n)^2) / (2 rows) and dL/dparams (float64).

    plant (tests only) introduces a known mistake so that tests can show
    their checks catch it: "transpose_w2" back-propagates through W2
    instead of W2^T, "drop_mask1" forgets the first ReLU's mask."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params)
    out, (xb, h1, h1b, h2, h2b) = forward(params, norm, x, bf)
    _, yn = _normalise(x, y, norm, bf)
    rows = len(out)
    e = out - yn
    loss = float(np.sum(e * e) / (2.0 * rows))
    d = e / rows
    db = r(d)
    gW3 = h2b.T @ db
    gb3 = d.sum(axis=0)
    dz2 = (db @ r(W3).T) * (h2 > 0)
    dz2b = r(dz2)
    gW2 = h1b.T @ dz2b
    gb2 = dz2.sum(axis=0)
    back2 = r(W2) if plant == "transpose_w2" else r(W2).T
    mask1 = 1.0 if plant == "drop_mask1" else (h1 > 0)
    dz1 = (dz2b @ back2) * mask1
    gW1 = xb.T @ r(dz1)
    gb1 = dz1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def init_params(seed: int) -> np.ndarray:
    """The library's fresh model (kcl_cost_model_create): a splitmix64 stream,
    u = (z >> 11) / 2^53, weight = (2u - 1) / sqrt(fan_in), zero biases."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF

    def nxt():
        nonlocal state
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    p = np.zeros(P, np.float32)
    for off, count, fan in ((OFF_W1, F * H, F), (OFF_W2, H * H, H), (OFF_W3, H * O, H)):
        bound = np.float32(1.0) / np.sqrt(np.float32(fan))
        for i in range(count):
            u = (nxt() >> 11) * (1.0 / 9007199254740992.0)
            p[off + i] = np.float32((2.0 * u - 1.0) * float(bound))
    return p
