"""CPU reference of the estimator network (MLP 256-128-64-3, ReLU, MSE).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the GPU forward pass and
gradients.  `bf16=True` rounds the GEMM operands to bfloat16
(round-to-nearest-even) at the same points as the GPU kernels, with float64
accumulation; `bf16=False` is the plain float64 network used to pin the
analytic backward pass against finite differences.
"""
from __future__ import annotations

import numpy as np

DIMS = [(256, 128), (128, 64), (64, 3)]


def bf16(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    """Flat parameter vector -> [(Wt, b)] in the C ABI order."""
    out, i = [], 0
    for fin, fout in DIMS:
        w = params[i:i + fin * fout].reshape(fout, fin)
        i += fin * fout
        b = params[i:i + fout]
        i += fout
        out.append((w.astype(np.float64), b.astype(np.float64)))
    assert i == len(params)
    return out


def forward(params, x, bf16_ops: bool = True):
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params)
    acts, h = [], r(x.astype(np.float32))
    for li, (w, b) in enumerate(layers):
        z = h @ r(w.astype(np.float32)).T + b
        if li < len(layers) - 1:
            z = np.maximum(z, 0.0)
        z = z.astype(np.float32).astype(np.float64) if bf16_ops else z
        acts.append((h, z))
        h = r(z.astype(np.float32)) if bf16_ops else z
    return acts[-1][1], acts


def loss_and_grads(params, x, t, bf16_ops: bool = True):
    """Mean over rows of the summed squared error, and its gradients."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    n = len(x)
    y, acts = forward(params, x, bf16_ops)
    loss = float(((y - t) ** 2).sum() / n)
    layers = unpack(params)
    grads = [None] * len(layers)
    dz = 2.0 * (y - t) / n
    if bf16_ops:
        dz = dz.astype(np.float32).astype(np.float64)
    for li in range(len(layers) - 1, -1, -1):
        h_in, _ = acts[li]
        w, _ = layers[li]
        gw = r(dz.astype(np.float32)).T @ h_in
        gb = dz.sum(axis=0)
        grads[li] = (gw, gb)
        if li > 0:
            dx = r(dz.astype(np.float32)) @ r(w.astype(np.float32))
            prev = acts[li - 1][1]
            dz = np.where(prev > 0, dx, 0.0)
            if bf16_ops:
                dz = dz.astype(np.float32).astype(np.float64)
    flat = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, flat
