"""CPU oracle of the canopy surrogate MLP (cable_amd/csrc/mlp.hip).

Test infrastructure only: tests/ use it to check the GPU forward pass, the
backward pass and Adam.  It restates the network in numpy:

    z1 = x.W1 + b1, h1 = hswish(z1), ..., y = h3.W4 + b4     (24-128-128-128-8)
    loss = mean((y - t)^2)

and its gradient by the chain rule.  With ``bf16=True`` every operand of a
matrix product is rounded to BF16 first, as the matrix cores see it, and the
products accumulate in FP64; with ``bf16=False`` the whole network is FP64,
which is what the finite-difference check in tests/test_mlp.py pins.
Parameter packing follows include/cable_mlp.h.
"""
from __future__ import annotations

import numpy as np

NIN, HID, NOUT = 24, 128, 8
SHAPES = [(NIN, HID), (HID, HID), (HID, HID), (HID, NOUT)]
NPARAM = sum(i * o + o for i, o in SHAPES)


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    Ws, bs, o = [], [], 0
    for i, n in SHAPES:
        Ws.append(p[o:o + i * n].reshape(i, n))
        o += i * n
        bs.append(p[o:o + n])
        o += n
    assert o == NPARAM
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in zip(Ws, bs)])


def hswish(z):
    """Hard-swish, z * min(max(z + 3, 0), 6) / 6 (Howard et al., MobileNetV3)."""
    return z * np.clip(z + 3.0, 0.0, 6.0) / 6.0


def hswish_grad(z):
    return np.where(z < -3.0, 0.0, np.where(z > 3.0, 1.0, (2.0 * z + 3.0) / 6.0))


def _mm(a, b, use_bf16):
    if use_bf16:
        return bf16(a) @ bf16(b)
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


def forward(params, x, bf16_ops=True):
    Ws, bs = unpack(params)
    h = np.asarray(x, dtype=np.float64)
    zs, hs = [], [h]
    for l in range(4):
        z = _mm(h, Ws[l], bf16_ops) + bs[l]
        if l < 3:
            zs.append(z)
            h = hswish(z)
            hs.append(h)
        else:
            h = z
    return h, (zs, hs)


def loss_and_grad(params, x, t, bf16_ops=True):
    """Mean-squared error over all n*8 outputs and its gradient (packed)."""
    Ws, bs = unpack(params)
    y, (zs, hs) = forward(params, x, bf16_ops)
    n = y.shape[0]
    d = y - t
    loss = float(np.mean(d * d))
    dcur = 2.0 * d / (n * NOUT)
    gW, gb = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        gW[l] = _mm(hs[l].T, dcur, bf16_ops)
        gb[l] = dcur.sum(axis=0)
        if l == 0:
            break
        dcur = _mm(dcur, Ws[l].T, bf16_ops) * hswish_grad(zs[l - 1])
    return loss, pack(gW, gb)


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new params, m, v."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def init_params(seed=0):
    """He-uniform weights for the hard-swish layers, zero biases, small output layer."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for l, (i, o) in enumerate(SHAPES):
        lim = np.sqrt(6.0 / i) * (0.5 if l == 3 else 1.0)
        Ws.append(rng.uniform(-lim, lim, (i, o)))
        bs.append(np.zeros(o))
    return pack(Ws, bs)
