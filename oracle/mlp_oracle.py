"""NumPy restatement of the MLP surrogate (forward, MSE loss, gradients).

TEST INFRASTRUCTURE ONLY (checker of nosh_amd's MFMA kernels; imported by
tests/ only). Parameters use the library's layout: all weights layer by layer,
each [out][in] row-major, then all biases. ``bf16=True`` rounds weights,
hidden activations and back-propagated pre-activation gradients to bfloat16
(round to nearest even) where the GPU kernels store them in BF16, leaving
only accumulation order and transcendental ulps as differences.
"""
from __future__ import annotations

import numpy as np

_C = np.sqrt(2.0 / np.pi)


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + 0.044715 * z ** 3)))


def dgelu(z):
    t = np.tanh(_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * z * z)


def dims_of(nin, width, depth, nout):
    return [nin] + [width] * depth + [nout]


def unpack(p, dims):
    Ws, bs, off = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(p[off:off + a * b].reshape(b, a))
        off += a * b
    for b in dims[1:]:
        bs.append(p[off:off + b])
        off += b
    return Ws, bs


def forward(p, X, dims, bf=False):
    """Returns (y, zs, hs): pre-activations and activations per hidden layer."""
    Ws, bs = unpack(np.asarray(p, np.float64), dims)
    if bf:
        Ws = [bf16(W) for W in Ws]
    h = np.asarray(X, np.float64)
    zs, hs = [], []
    for l in range(len(Ws) - 1):
        z = h @ Ws[l].T + bs[l]
        h = gelu(z)
        if bf:
            h = bf16(h)
        zs.append(z)
        hs.append(h)
    y = h @ Ws[-1].T + bs[-1]
    return y, zs, hs


def loss_and_grad(p, X, T, dims, bf=False):
    """Mean squared error over all outputs and its gradient (parameter layout)."""
    X = np.asarray(X, np.float64)
    Ws, _ = unpack(np.asarray(p, np.float64), dims)
    if bf:
        Ws = [bf16(W) for W in Ws]
    y, zs, hs = forward(p, X, dims, bf)
    n = y.size
    d = y - T
    loss = float(np.sum(d * d) / n)
    dy = 2.0 * d / n
    gW = [None] * len(Ws)
    gb = [None] * len(Ws)
    L = len(Ws)
    gW[L - 1] = dy.T @ hs[-1]
    gb[L - 1] = dy.sum(axis=0)
    dz = (dy @ Ws[L - 1]) * dgelu(zs[-1])
    if bf:
        dz = bf16(dz)
    for l in range(L - 2, 0, -1):
        gW[l] = dz.T @ hs[l - 1]
        gb[l] = dz.sum(axis=0)
        dz = (dz @ Ws[l]) * dgelu(zs[l - 1])
        if bf:
            dz = bf16(dz)
    gW[0] = dz.T @ X
    gb[0] = dz.sum(axis=0)
    g = np.concatenate([w.ravel() for w in gW] + [b for b in gb])
    return loss, g
