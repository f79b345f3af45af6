"""CPU oracle of the branch-scorer MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the HIP scorer kernels
(traccc_amd/csrc/scorer/mlp.hip); the product never calls it.

The network: 12 -> 64 -> 64 -> 1, GELU (exact, erf form), on features
normalised with a per-feature mean and standard deviation; loss: mean
binary cross-entropy with logits, survive labels weighted by pos_weight.
Parameters are one flat vector W1[12][64] b1[64] W2[64][64] b2[64] W3[64] b3.

Two modes:
  * bf16=False: FP64 throughout (pinned by finite differences and closed
    forms in tests/test_scorer.py);
  * bf16=True: the same FP64 arithmetic with values rounded to BF16 exactly
    where the device kernels hand BF16 operands to the matrix cores (inputs,
    W1, W2, W3, H1, and in the backward pass dZ2, dZ1), so the device and the
    oracle differ only by FP32 accumulation order.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN, N_HID = 12, 64
N_PARAMS = N_IN * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID + 1
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_IN, N_HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _q(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(p, X, mean, std, bf16_points=False):
    """Logits and the intermediates the backward pass uses."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    inv = 1.0 / np.where(np.asarray(std, np.float64) > 1e-12, std, 1.0)
    Xn = _q((np.asarray(X, np.float64) - mean) * np.asarray(inv, np.float32), bf16_points)
    Z1 = Xn @ _q(W1, bf16_points) + b1
    H1 = _q(gelu(Z1), bf16_points)
    Z2 = H1 @ _q(W2, bf16_points) + b2
    H2 = gelu(Z2)
    s = H2 @ _q(W3, bf16_points) + b3
    return s, (Xn, Z1, H1, Z2, H2)


def loss_and_grad(p, X, y, mean, std, pos_weight=1.0, bf16_points=False):
    """Mean weighted BCE with logits and its gradient (flat, FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    s, (Xn, Z1, H1, Z2, H2) = forward(p, X, mean, std, bf16_points)
    w = np.where(y > 0.5, pos_weight, 1.0) / n
    loss = np.sum(w * (np.maximum(s, 0) - s * y + np.log1p(np.exp(-np.abs(s)))))
    dS = w * (1.0 / (1.0 + np.exp(-s)) - y)
    g = np.zeros(N_PARAMS)
    W3q = _q(W3, bf16_points)
    g[OFF_B3] = dS.sum()
    g[OFF_W3:OFF_B3] = H2.T @ dS
    dZ2 = np.outer(dS, W3q) * gelu_grad(Z2)
    g[OFF_B2:OFF_W3] = dZ2.sum(axis=0)
    dZ2q = _q(dZ2, bf16_points)
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2q).ravel()
    dH1 = dZ2q @ _q(W2, bf16_points).T
    dZ1 = dH1 * gelu_grad(Z1)
    g[OFF_B1:OFF_W2] = dZ1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (Xn.T @ _q(dZ1, bf16_points)).ravel()
    return loss, g
