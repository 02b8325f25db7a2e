"""CPU restatement of the closure network (test infrastructure only).

The network has no counterpart in the reference; this module is the CPU
check of the MFMA kernels in peridem_amd/csrc/pdem_mlp.hpp.  It evaluates
the same 16 -> 128 -> 128 -> 128 -> 4 ReLU perceptron in float64, with the
kernels' BF16 roundings applied at the same points (inputs after
normalisation, weights, hidden activations, output gradient and the
back-propagated pre-activation gradients) when ``bf16=True``.  With
``bf16=False`` it is a plain float64 network whose gradient is pinned by
central finite differences (tests/test_mlp_oracle.py).

Parameter vector layout (row-major W[in][out]):
W1[16][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][4] b4[4].
Loss: mean over rows of sum_o w_o (y_o - t_o s_o)^2 with y the normalised
output, t the target, s the output scale and w the output weight.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 16, 128, 4
SHAPES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(q, dtype=np.float64).ravel() for q in parts])


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def forward(p, x, in_mean, in_scale, bf16_on=True):
    """Returns (y_normalised, cache)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float32)
    if bf16_on:
        xn = bf16((x - np.asarray(in_mean, np.float32)) * np.asarray(in_scale, np.float32))
    else:
        xn = (x.astype(np.float64) - np.asarray(in_mean, np.float64)) * np.asarray(in_scale, np.float64)
    Wb = [_r(W, bf16_on) for W in (W1, W2, W3, W4)]
    z1 = xn @ Wb[0] + b1
    h1 = _r(np.maximum(z1, 0.0), bf16_on)
    z2 = h1 @ Wb[1] + b2
    h2 = _r(np.maximum(z2, 0.0), bf16_on)
    z3 = h2 @ Wb[2] + b3
    h3 = _r(np.maximum(z3, 0.0), bf16_on)
    y = h3 @ Wb[3] + b4
    return y, (xn, h1, h2, h3, Wb)


def loss_and_grad(p, x, t, in_mean, in_scale, out_scale, out_weight, bf16_on=True):
    """Loss and its gradient with respect to the parameter vector."""
    y, (xn, h1, h2, h3, Wb) = forward(p, x, in_mean, in_scale, bf16_on)
    n = x.shape[0]
    tn = np.asarray(t, np.float32) * np.asarray(out_scale, np.float32)
    e = y - tn.astype(np.float64)
    loss = float((out_weight * e * e).sum() / n)
    dy = _r(2.0 * out_weight * e / n, bf16_on)
    gW4 = h3.T @ dy
    gb4 = dy.sum(0)
    dz3 = np.where(h3 > 0, _r(dy @ Wb[3].T, bf16_on), 0.0)
    gW3 = h2.T @ dz3
    gb3 = dz3.sum(0)
    dz2 = np.where(h2 > 0, _r(dz3 @ Wb[2].T, bf16_on), 0.0)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = np.where(h1 > 0, _r(dz2 @ Wb[1].T, bf16_on), 0.0)
    gW1 = xn.T @ dz1
    gb1 = dz1.sum(0)
    return loss, pack([gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4])


def predict(p, x, in_mean, in_scale, out_scale, bf16_on=True):
    y, _ = forward(p, x, in_mean, in_scale, bf16_on)
    return y / np.asarray(out_scale, np.float64)
