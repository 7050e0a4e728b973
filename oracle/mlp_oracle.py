"""Test infrastructure, not product: a numpy statement of the branching-score
MLP (in -> h1 -> h2 -> 1, ReLU) that the GPU kernels of
sovereign_amd/csrc/surrogate.hip implement, used only by tests/ to check
them. The network and its training have no counterpart in the reference
(whose pick_strong_branch, solver/milp/src/branch_and_bound.cpp:726-793,
ranks candidates by fractionality); what is restated here is the
mathematics of the kernels, with their rounding points:

  every matrix product rounds both operands to BF16 (round to nearest even)
  and accumulates (here in FP64, on the GPU in FP32);
  Z1 = X W1 + b1, H1 = relu(Z1); Z2 = H1 W2 + b2, H2 = relu(Z2);
  y = H2 w3 + b3; loss = 0.5 mean (y - t)^2;
  dy = (y - t) / B; gW3 = H2^T dy; gb3 = sum dy;
  dH2 = (dy w3^T) * [H2 > 0]; gW2 = H1^T dH2; gb2 = colsum dH2;
  dH1 = (dH2 W2^T) * [H1 > 0]; gW1 = X^T dH1; gb1 = colsum dH1.

Parameters are packed as W1 (in x h1), b1, W2 (h1 x h2), b2, w3 (h2), b3,
row-major, as in include/sovereign_surrogate.h. With round=False the same
functions are a plain FP64 network; its gradients are pinned against finite
differences of its loss (tests/test_surrogate.py)."""
from __future__ import annotations

import numpy as np


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray, shape):
    i, h1, h2 = shape
    o = 0
    W1 = p[o:o + i * h1].reshape(i, h1); o += i * h1
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h1 * h2].reshape(h1, h2); o += h1 * h2
    b2 = p[o:o + h2]; o += h2
    w3 = p[o:o + h2]; o += h2
    b3 = p[o:o + 1]; o += 1
    assert o == len(p)
    return W1, b1, W2, b2, w3, b3


def param_count(shape) -> int:
    i, h1, h2 = shape
    return i * h1 + h1 + h1 * h2 + h2 + h2 + 1


def _mm(a, b, rnd):
    if rnd:
        return bf16(a) @ bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(p, shape, X, round=True):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64), shape)
    X = np.asarray(X, np.float64)
    H1 = np.maximum(_mm(X, W1, round) + b1, 0.0)
    H2 = np.maximum(_mm(H1, W2, round) + b2, 0.0)
    y = _mm(H2, w3[:, None], round)[:, 0] + b3[0]
    if round:  # the GPU keeps activations in FP32 between the products
        H1, H2, y = (v.astype(np.float32).astype(np.float64) for v in (H1, H2, y))
    return H1, H2, y


def loss_and_gradients(p, shape, X, t, round=True):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64), shape)
    X = np.asarray(X, np.float64)
    t = np.asarray(t, np.float64)
    H1, H2, y = forward(p, shape, X, round)
    B = len(t)
    d = y - t
    loss = 0.5 * np.mean(d * d)
    dy = d / B
    gW3 = _mm(H2.T, dy[:, None], round)[:, 0]
    gb3 = np.array([dy.sum()])
    dH2 = _mm(dy[:, None], w3[None, :], round) * (H2 > 0)
    gW2 = _mm(H1.T, dH2, round)
    gb2 = dH2.sum(axis=0)
    dH1 = _mm(dH2, W2.T, round) * (H1 > 0)
    gW1 = _mm(X.T, dH1, round)
    gb1 = dH1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g


def he_init(shape, seed):
    """Any He-normal draw (the oracle checks products, not the GPU's initial draw)."""
    rng = np.random.default_rng(seed)
    i, h1, h2 = shape
    return np.concatenate([rng.normal(0, np.sqrt(2 / i), i * h1), np.zeros(h1),
                           rng.normal(0, np.sqrt(2 / h1), h1 * h2), np.zeros(h2),
                           rng.normal(0, np.sqrt(2 / h2), h2), np.zeros(1)])
