"""NumPy restatement of the importance-sampling MLP — TEST INFRASTRUCTURE ONLY.

The checker of the GPU forward/backward kernels (cusmc_amd/csrc/is_kernels.hip,
is.cpp).  No counterpart in the reference: cuSMC has no learned proposal.

Network: theta = clamp(W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3) per
location, zero at untiltable locations.  The GPU rounds GEMM inputs (x, the
weights, the hidden activations and the backward signals) to bf16 and
accumulates in fp32; `forward`/`backward` below do the same roundings in
float64 arithmetic, so the two agree to fp32 accumulation-order error.

Parameter layout (flat fp32): W1 [fpad x H], b1 [H], W2 [H x H], b2 [H],
W3 [H x gpad], b3 [gpad] — each W row-major (inputs x outputs).  The network
has one output per tilt group; location l reads output group[l] (`expand`),
and the gradient of a group output is the sum over its locations
(`group_sum`).
"""
from __future__ import annotations

import numpy as np


def bf16_round(a) -> np.ndarray:
    """Round to the nearest bf16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(x.shape)


def bf16_bits_to_f32(bits) -> np.ndarray:
    b = np.ascontiguousarray(bits, dtype=np.uint16)
    return (b.astype(np.uint32) << 16).view(np.float32).reshape(b.shape)


def unpack(params: np.ndarray, fpad: int, H: int, lpad: int, dtype=np.float32):
    p = np.asarray(params, dtype=dtype)
    o = 0
    out = []
    for shape in ((fpad, H), (H,), (H, H), (H,), (H, lpad), (lpad,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size, (o, p.size)
    return out


def forward(params, x, fpad, H, lpad, tilt_max, tiltable):
    """Returns theta and the cache of the backward pass; x: [rows, fpad]."""
    W1, b1, W2, b2, W3, b3 = unpack(params, fpad, H, lpad)
    xb = bf16_round(x).astype(np.float64)
    z1 = xb @ bf16_round(W1).astype(np.float64) + b1
    h1 = bf16_round(np.maximum(z1, 0.0)).astype(np.float64)
    z2 = h1 @ bf16_round(W2).astype(np.float64) + b2
    h2 = bf16_round(np.maximum(z2, 0.0)).astype(np.float64)
    z3 = h2 @ bf16_round(W3).astype(np.float64) + b3
    theta = np.clip(z3, -tilt_max, tilt_max)
    theta[:, ~np.asarray(tiltable, dtype=bool)] = 0.0
    return theta, (xb, h1, h2, W2, W3)


def backward(params, cache, g, fpad, H, lpad):
    """Gradient of sum(g * theta) over the flat parameters, given g already
    zero at clamped and untiltable outputs (the GPU's convention)."""
    xb, h1, h2, W2, W3 = cache
    gb = bf16_round(g).astype(np.float64)
    gW3 = h2.T @ gb
    gb3 = np.asarray(g, dtype=np.float64).sum(0)
    dh2 = (gb @ bf16_round(W3).astype(np.float64).T) * (h2 > 0)
    dh2b = bf16_round(dh2).astype(np.float64)
    gW2 = h1.T @ dh2b
    gb2 = dh2.sum(0)
    dh1 = (dh2b @ bf16_round(W2).astype(np.float64).T) * (h1 > 0)
    dh1b = bf16_round(dh1).astype(np.float64)
    gW1 = xb.T @ dh1b
    gb1 = dh1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])


def forward_exact(params, x, fpad, H, lpad, tilt_max, tiltable):
    """The same network in float64 without bf16 rounding (for finite differences)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, fpad, H, lpad, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    z3 = h2 @ W3 + b3
    theta = np.clip(z3, -tilt_max, tilt_max)
    theta[:, ~np.asarray(tiltable, dtype=bool)] = 0.0
    return theta


def backward_exact(params, x, g, fpad, H, lpad, tilt_max, tiltable):
    """Analytic gradient of sum(g * theta) in float64 without rounding."""
    W1, b1, W2, b2, W3, b3 = unpack(params, fpad, H, lpad, np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ W3 + b3
    gz3 = np.array(g, dtype=np.float64)
    gz3[np.abs(z3) >= tilt_max] = 0.0
    gz3[:, ~np.asarray(tiltable, dtype=bool)] = 0.0
    gW3 = h2.T @ gz3
    gb3 = gz3.sum(0)
    dh2 = (gz3 @ W3.T) * (z2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])


def ce_output_grad(theta, A, B, w, tiltable, tilt_max):
    """d(-sum_pairs w ln q)/d theta: -w ln2 (A - B 2^theta), zero where clamped
    or untiltable (the statistics A, B of each recorded macro-step)."""
    th = np.asarray(theta, dtype=np.float64)
    g = -np.asarray(w, dtype=np.float64)[:, None] * np.log(2.0) * (A - B * np.exp2(th))
    g[np.abs(th) >= tilt_max] = 0.0
    g[:, ~np.asarray(tiltable, dtype=bool)] = 0.0
    return g


def expand(theta_group, group, lpad):
    """theta per location (padding locations 0)."""
    tg = np.asarray(theta_group)
    out = np.zeros((tg.shape[0], lpad), dtype=tg.dtype)
    out[:, :len(group)] = tg[:, np.asarray(group)]
    return out


def group_sum(g_loc, group, gpad):
    """Gradient of each group output: sum over its member locations."""
    g = np.asarray(g_loc, dtype=np.float64)
    out = np.zeros((g.shape[0], gpad))
    np.add.at(out.T, np.asarray(group), g[:, :len(group)].T)
    return out
