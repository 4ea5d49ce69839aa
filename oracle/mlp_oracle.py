"""CPU oracle of the learned stress model (TEST INFRASTRUCTURE ONLY).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark may
use this module; the engine never does.

The reference has no learned closure, so there is no reference file to
restate: this is the textbook arithmetic of a 6-64-64-3 ReLU perceptron
(forward pass, backpropagation of the mean squared error, Adam), in FP64,
with optional rounding to BF16 at the points where the HIP kernels round
(squishy_volumes_amd/csrc/closure.hip): weights, inputs, hidden activations
and back-propagated errors where they enter a product.  Values are first
rounded to FP32 (the kernels compute in FP32) and then to BF16, nearest even.

Parity pins: the gradient is checked against central finite differences of
the unrounded loss (tests/test_closure.py), and bf16() against torch's
float32 -> bfloat16 conversion.

Parameters use the engine's padded layout (include/mpm_engine.h):
W1 [64][16], b1 [64], W2 [64][64], b2 [64], W3 [16][64], b3 [16].
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT, K_IN, K_OUT = 6, 64, 3, 16, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * K_IN
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + K_OUT * N_H
N_PARAM = OFF_B3 + K_OUT


def bf16(x) -> np.ndarray:
    """x rounded to FP32, then to the nearest BF16 (ties to even), as float64."""
    u = np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_H, K_IN)[:, :N_IN]
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_H, N_H)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(K_OUT, N_H)[:N_OUT]
    b3 = p[OFF_B3:][:N_OUT]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    p = np.zeros(N_PARAM, dtype=np.float64)
    p[OFF_W1:OFF_B1].reshape(N_H, K_IN)[:, :N_IN] = W1
    p[OFF_B1:OFF_W2] = b1
    p[OFF_W2:OFF_B2] = np.asarray(W2).ravel()
    p[OFF_B2:OFF_W3] = b2
    p[OFF_W3:OFF_B3].reshape(K_OUT, N_H)[:N_OUT] = W3
    p[OFF_B3:OFF_B3 + N_OUT] = b3
    return p


def forward(p, x, rounded: bool = True):
    """Outputs y [n][3] and the activations (h1, h2) of samples x [n][6]."""
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xb = r(x)
    h1 = r(np.maximum(xb @ r(W1).T + b1, 0.0))
    h2 = r(np.maximum(h1 @ r(W2).T + b2, 0.0))
    y = h2 @ r(W3).T + b3
    return y, (xb, h1, h2)


def loss_grad(p, x, t, scale: float, rounded: bool = True):
    """(sum of squared errors, scale * its gradient in the padded layout)."""
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (xb, h1, h2) = forward(p, x, rounded)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d))
    dy = r(2.0 * scale * d)
    d2 = np.where(h2 > 0, r(dy @ r(W3)), 0.0)
    d1 = np.where(h1 > 0, r(d2 @ r(W2)), 0.0)
    g = pack(d1.T @ xb, d1.sum(0), d2.T @ h1, d2.sum(0), dy.T @ h2, dy.sum(0))
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    """One bias-corrected Adam step; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1**step)) / (np.sqrt(v / (1 - b2**step)) + eps)
    return p, m, v
