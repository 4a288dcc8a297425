"""CPU oracle for the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product never does.  It restates, in
numpy, what libctl's MLP kernels compute (cutauleaping_amd/csrc/ctl_mlp.hip;
the surrogate has no counterpart in the reference):

* parameters: one flat FP32 vector W_1 [w0][w1], b_1 [w1], W_2, b_2, ...
* forward: h_0 = bf16(x); h_l = bf16(relu(h_{l-1} @ bf16(W_l) + b_l)) for the
  hidden layers; y = h_{L-1} @ bf16(W_L) + b_L in FP32;
* loss: mean of (y - t)^2 over all rows * outputs;
* backward: d_L = bf16(2 (y - t) / (B wL)); dW_l = h_{l-1}^T d_l;
  db_l = column sums of d_l; d_{l-1} = bf16((d_l @ bf16(W_l)^T) * [h_{l-1} > 0]).

Products accumulate in float64 here (FP32 with an unspecified order on the
GPU), so GPU/oracle agreement is to a stated tolerance, not bit-exact.

Pinning: with ``bf16=False`` and float64 everything, ``gradient`` is checked
against central finite differences of ``loss`` (tests/test_oracle_mlp.py),
which fixes the backward pass's signs, transposes and ReLU mask; ``bf16_round``
is checked against torch's bfloat16 conversion (round to nearest even).
"""
from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), returned as FP32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    quiet = np.where((u & 0xFFFF) != 0, np.uint64(0x400000), np.uint64(0))
    r = np.where(nan, (u & np.uint64(0xFFFF0000)) | quiet, r)
    return (r & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.float32).reshape(a.shape)


def layout(widths):
    """[(weight offset, bias offset)] per layer and the parameter count."""
    offs, o = [], 0
    for a, b in zip(widths[:-1], widths[1:]):
        offs.append((o, o + a * b))
        o += a * b + b
    return offs, o


def unpack(p, widths):
    offs, n = layout(widths)
    assert p.size == n
    Ws, bs = [], []
    for (wo, bo), a, b in zip(offs, widths[:-1], widths[1:]):
        Ws.append(p[wo:wo + a * b].reshape(a, b))
        bs.append(p[bo:bo + b])
    return Ws, bs


def _q(x, bf16):
    return bf16_round(x).astype(np.float64) if bf16 else np.asarray(x, dtype=np.float64)


def forward(p, widths, x, bf16=True, keep=False):
    Ws, bs = unpack(np.asarray(p), widths)
    h = _q(x, bf16)
    acts = [h]
    L = len(Ws)
    for l in range(L):
        z = h @ _q(Ws[l], bf16) + np.asarray(bs[l], np.float64)
        if l + 1 < L:
            if bf16:
                z = z.astype(np.float32)
            h = _q(np.maximum(z, 0.0), bf16)
            acts.append(h)
        else:
            y = z.astype(np.float32) if bf16 else z
    return (y, acts) if keep else y


def loss(p, widths, x, t, bf16=True) -> float:
    y = forward(p, widths, x, bf16)
    return float(np.mean((np.asarray(y, np.float64) - np.asarray(t, np.float64)) ** 2))


def gradient(p, widths, x, t, bf16=True):
    """(loss, flat gradient) as the GPU backward computes them."""
    Ws, bs = unpack(np.asarray(p), widths)
    y, acts = forward(p, widths, x, bf16, keep=True)
    B, wL = y.shape
    e = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    lval = float(np.mean(e * e))
    d = 2.0 * e / (B * wL)
    if bf16:
        d = _q(d.astype(np.float32), True)
    offs, n = layout(widths)
    g = np.zeros(n, np.float64)
    for l in range(len(Ws) - 1, -1, -1):
        wo, bo = offs[l]
        a, b = widths[l], widths[l + 1]
        g[wo:wo + a * b] = (acts[l].T @ d).ravel()
        g[bo:bo + b] = d.sum(axis=0)
        if l > 0:
            dp = (d @ _q(Ws[l], bf16).T) * (acts[l] > 0)
            d = _q(dp.astype(np.float32), True) if bf16 else dp
    return lval, g
