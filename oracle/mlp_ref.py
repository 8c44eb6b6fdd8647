"""TEST INFRASTRUCTURE ONLY — FP64 CPU reference of the warm-start network (forward, MSE loss, backward).

The network has no counterpart in the reference; this is the checker for the HIP kernels behind include/ifea_mlp.h.
Both operands of every product are rounded to BF16 (round to nearest even) before use, as the device's BF16 MFMA
products take them (weights, activations, and the backward's output gradients; the bias gradient is summed from the
unrounded gradients), so the comparison isolates the FP32-accumulation error of the kernels. With bf16=False it is
the exact network, pinned by finite differences in tests/test_mlp.py.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def to_bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params, shapes):
    out, o = [], 0
    for fi, fo in shapes:
        W = np.asarray(params[o:o + fi * fo], np.float64).reshape(fi, fo)
        o += fi * fo
        b = np.asarray(params[o:o + fo], np.float64)
        o += fo
        out.append((W, b))
    return out


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_d(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(params, shapes, x, bf16=True):
    layers = unpack(params, shapes)
    a = np.asarray(x, np.float64)
    zs, acts = [], [a]
    for l, (W, b) in enumerate(layers):
        Wq = to_bf16(W.astype(np.float32)) if bf16 else W
        z = (to_bf16(a) if bf16 else a) @ Wq + b
        zs.append(z)
        a = gelu(z) if l < len(layers) - 1 else z
        acts.append(a)
    return a, zs, acts


def loss_grad(params, shapes, x, y, bf16=True):
    """Mean squared error over all outputs and its gradient (same parameter layout as the device)."""
    layers = unpack(params, shapes)
    yh, zs, acts = forward(params, shapes, x, bf16)
    y = np.asarray(y, np.float64)
    n = yh.size
    loss = np.mean((yh - y) ** 2)
    dz = 2.0 * (yh - y) / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        Wq = to_bf16(W.astype(np.float32)) if bf16 else W
        q = to_bf16 if bf16 else (lambda v: v)
        gW = q(acts[l]).T @ q(dz)
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = (q(dz) @ Wq.T) * gelu_d(zs[l - 1])
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])
    return loss, flat
