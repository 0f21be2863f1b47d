"""FP64 CPU restatement of the score-surrogate network (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the GPU network in vship_amd/csrc/mlp.hip; nothing
in the product imports it. The network has no counterpart in the reference:
it follows the description in include/vship_amd.h (55 -> 128 -> 128 -> 2,
exact GELU, Gaussian negative log-likelihood). Weights are rounded to BF16
exactly as the GPU copies are (round to nearest even), then every operation is
float64, so the comparison measures the GPU's FP32 accumulation only.

Pins: tests/test_oracle.py checks loss_and_grad against central finite
differences of loss_and_grad's own loss (a wrong sign, index or transposed
weight fails it), GELU against its closed-form values and derivative, and the
loss against its value at the optimum.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, HID, OUT = 55, 128, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + OUT * HID
PARAMS = OFF_B3 + OUT


def bf16_round(a) -> np.ndarray:
    """float32 values rounded to the nearest BF16 (ties to even), as float32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params, quantise: bool = True):
    """(W1, b1, W2, b2, W3, b3) in float64; weight matrices BF16-rounded."""
    p = np.asarray(params, dtype=np.float32)
    assert p.size == PARAMS
    q = bf16_round if quantise else (lambda x: x)
    W1 = q(p[OFF_W1:OFF_B1]).astype(np.float64).reshape(HID, IN)
    b1 = p[OFF_B1:OFF_W2].astype(np.float64)
    W2 = q(p[OFF_W2:OFF_B2]).astype(np.float64).reshape(HID, HID)
    b2 = p[OFF_B2:OFF_W3].astype(np.float64)
    W3 = q(p[OFF_W3:OFF_B3]).astype(np.float64).reshape(OUT, HID)
    b3 = p[OFF_B3:].astype(np.float64)
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def normalise(X, xmean, xstd):
    """(X - mean) * inv with inv = float32(1 / std), 0 for std <= 0 (the GPU's
    normalisation constants)."""
    xstd = np.asarray(xstd, dtype=np.float32)
    inv = np.where(xstd > 0, np.float32(1.0) / np.where(xstd > 0, xstd, np.float32(1.0)), np.float32(0.0))
    return (np.asarray(X, np.float64) - np.asarray(xmean, np.float32).astype(np.float64)) * inv.astype(np.float64)


def forward(params, Xn, quantise: bool = True):
    W1, b1, W2, b2, W3, b3 = unpack(params, quantise)
    z1 = Xn @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    out = h2 @ W3.T + b3
    return out, (z1, h1, z2, h2)


def loss_and_grad(params, X, Y, xmean, xstd, ymean, ystd, quantise: bool = True):
    """Mean Gaussian NLL 0.5 * (lv + (y - mu)^2 exp(-lv)) over the samples, on
    the normalised target, and its gradient with respect to the parameter
    vector (through the BF16-rounded weights: d/dW of the rounded value)."""
    Xn = normalise(X, xmean, xstd)
    yn = (np.asarray(Y, np.float64) - np.float64(np.float32(ymean))) * np.float64(np.float32(1.0) / np.float32(ystd))
    n = Xn.shape[0]
    out, (z1, h1, z2, h2) = forward(params, Xn, quantise)
    W1, b1, W2, b2, W3, b3 = unpack(params, quantise)
    mu, lv = out[:, 0], out[:, 1]
    r = yn - mu
    e = np.exp(-lv)
    loss = float(np.mean(0.5 * (lv + r * r * e)))
    dout = np.stack([-r * e, 0.5 * (1.0 - r * r * e)], axis=1) / n
    g = np.zeros(PARAMS, dtype=np.float64)
    g[OFF_W3:OFF_B3] = (dout.T @ h2).ravel()
    g[OFF_B3:] = dout.sum(0)
    dz2 = (dout @ W3) * gelu_grad(z2)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    g[OFF_W1:OFF_B1] = (dz1.T @ Xn).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def predict(params, X, xmean, xstd, ymean, ystd):
    """Score mean and standard deviation in score units."""
    out, _ = forward(params, normalise(X, xmean, xstd))
    return out[:, 0] * ystd + ymean, np.exp(0.5 * out[:, 1]) * ystd
