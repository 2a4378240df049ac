"""FP64 CPU oracle of the long-wave emulator network -- TEST INFRASTRUCTURE.

Forward and backward of the 43 -> 256 -> 256 -> 256 -> 21 GELU network in
numpy float64, rounding to bfloat16 exactly where the GPU kernels
(exoplasim_amd/csrc/rad_mlp.hip) hold bf16 values: the layer inputs, the
weights, the activations and the back-propagated pre-activation gradients.
Sums are exact-order-independent FP64, so the comparison measures the
kernels' FP32 accumulation plus any bf16 rounding flips.

GELU is the erf form; its derivative is Phi(z) + z phi(z).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(x: np.ndarray) -> np.ndarray:
    """round to nearest even bfloat16, returned as float64"""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def exact(x: np.ndarray) -> np.ndarray:
    """no rounding: the plain FP64 network (for checking the calculus)"""
    return np.asarray(x, dtype=np.float64)


def forward(x, params, rnd=bf16):
    """x (M, 43) float; params [W1, b1, ..., W4, b4] with W (out, in).
    Returns y (M, 21) and the saved (inputs, pre-activations).  rnd is the
    rounding applied where the kernels hold bf16 values."""
    h = rnd(x)
    saved = []
    for li in range(4):
        w, b = params[2 * li], params[2 * li + 1]
        z = h @ rnd(w).T + b
        saved.append((h, z))
        h = rnd(gelu(z)) if li < 3 else z
    return h, saved


def backward(gy, params, saved, rnd=bf16):
    """gradients of sum(gy * y) w.r.t. W1, b1, ..., W4, b4"""
    grads = [None] * 8
    dz32 = gy.astype(np.float32).astype(np.float64) if rnd is bf16 else exact(gy)
    dz = rnd(gy)
    for li in (3, 2, 1, 0):
        h, z = saved[li]
        grads[2 * li] = dz.T @ h
        grads[2 * li + 1] = dz32.sum(axis=0)
        if li > 0:
            dh = dz @ rnd(params[2 * li])
            zp = saved[li - 1][1]
            g = dh * gelu_d(zp)
            dz32 = g
            dz = rnd(g)
    return grads
