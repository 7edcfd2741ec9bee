"""FP64 numpy restatement of the learned proposal's network.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
forward, backward and BF16-MFMA kernels; the product never calls it.

The reference has no learned proposal (its importance map is the retired
centroid-propagation pass, include/kernels/importanceSampling.hpp:10-18), so
nothing here follows a reference file: the network is the one fixed by the
C ABI (include/hase_ase.h, "Learned proposal"): 10 inputs -> 4 hidden SiLU
layers of 128 -> 1 output, parameters W1 b1 W2 b2 W3 b3 W4 b4 w5 b5 with
every W row-major [out][in].  Pinned by finite differences and closed-form
networks in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN = 10
H = 128
SIZES = [(H, IN), (H,), (H, H), (H,), (H, H), (H,), (H, H), (H,), (H,), (1,)]
NAMES = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "w5", "b5"]
COUNT = int(sum(np.prod(s) for s in SIZES))


def unpack(params):
    """Views of the flat parameter vector, in ABI order."""
    p = np.asarray(params)
    out, at = {}, 0
    for name, shape in zip(NAMES, SIZES):
        n = int(np.prod(shape))
        out[name] = p[at:at + n].reshape(shape)
        at += n
    assert at == p.size == COUNT
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[n], dtype=np.float64).reshape(-1) for n in NAMES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, *, keep=False):
    """y = net(x) for rows x (n, 10), FP64."""
    P = unpack(np.asarray(params, dtype=np.float64))
    a = np.asarray(x, dtype=np.float64)
    zs, acts = [], [a]
    for l in range(4):
        z = a @ P[f"W{l + 1}"].T + P[f"b{l + 1}"]
        a = silu(z)
        zs.append(z)
        acts.append(a)
    y = a @ P["w5"] + P["b5"][0]
    return (y, zs, acts) if keep else y


def loss_gradient(params, x, t, w=None):
    """Weighted squared error sum w (y - t)^2 / sum w and its FP64 gradient."""
    P = unpack(np.asarray(params, dtype=np.float64))
    t = np.asarray(t, dtype=np.float64)
    w = np.ones_like(t) if w is None else np.asarray(w, dtype=np.float64)
    W = w.sum()
    y, zs, acts = forward(params, x, keep=True)
    r = y - t
    loss = float((w * r * r).sum() / W)
    dy = 2.0 * w * r / W
    g = {"w5": acts[4].T @ dy, "b5": np.array([dy.sum()])}
    dz = np.outer(dy, P["w5"]) * silu_grad(zs[3])
    for l in range(3, -1, -1):
        g[f"W{l + 1}"] = dz.T @ acts[l]
        g[f"b{l + 1}"] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ P[f"W{l + 1}"]) * silu_grad(zs[l - 1])
    return loss, pack(g)


def to_bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(params, x):
    """Arithmetic of the MFMA inference kernel: BF16 weights and layer inputs,
    FP32 biases, FP64 stand-in for the FP32 accumulators, FP32 output layer on
    unrounded activations."""
    P = unpack(np.asarray(params, dtype=np.float32).astype(np.float64))
    a = to_bf16(x)
    z = None
    for l in range(4):
        z = a @ to_bf16(P[f"W{l + 1}"]).T + P[f"b{l + 1}"]
        a = to_bf16(silu(z))
    return silu(z) @ P["w5"] + P["b5"][0]


def init_params(seed=0):
    """He-style initialisation (uniform, variance 2/fan_in), zero biases."""
    rng = np.random.default_rng(seed)
    parts = {}
    for name, shape in zip(NAMES, SIZES):
        if name.startswith("b"):
            parts[name] = np.zeros(shape)
        else:
            fan_in = shape[-1] if len(shape) == 2 else H
            lim = np.sqrt(6.0 / fan_in)
            parts[name] = rng.uniform(-lim, lim, size=shape)
    return pack(parts)


def proposal_inputs(sx, sy, sz, cx, cy, level, beta, h, log_sigma):
    """Unnormalised network inputs (float32 arithmetic, as on the device)."""
    f32 = np.float32
    sx, sy, sz, cx, cy = (np.asarray(v, dtype=f32) for v in (sx, sy, sz, cx, cy))
    level = np.asarray(level, dtype=f32)
    beta = np.asarray(beta, dtype=f32)
    cz = (level + f32(0.5)) * f32(h)
    dx, dy, dz = sx - cx, sy - cy, sz - cz
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    n = np.broadcast(sx, sy, sz, cx, cy, level, beta).shape
    cols = [sx, sy, sz, cx, cy, cz, level, beta, dist, np.full(n, f32(log_sigma))]
    return np.stack([np.broadcast_to(c, n) for c in cols], axis=-1).astype(f32)
