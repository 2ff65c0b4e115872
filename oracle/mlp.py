"""CPU oracle of the proposal network -- TEST INFRASTRUCTURE ONLY.

The 2 -> 128 -> 128 -> 1 SiLU network of goofit_amd/csrc/gf_mlp.hip in numpy:
forward, the gradient of sum_i coef_i f(x_i), and the maximum-likelihood loss.
With ``bf16=True`` the operands of the 128 x 128 layer (h1, W2) and of both of
its backward products (delta2) are rounded to bfloat16 (round to nearest even)
exactly where the device kernels round them, and everything else is float64;
with ``bf16=False`` it is plain float64 (used to pin the analytic gradient
against finite differences).  No reference counterpart: GooFit has no learned
normalisation.
"""
from __future__ import annotations

import numpy as np

H = 128
W1, B1 = 0, 2 * H
W2 = 3 * H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
NPARAM = B3 + 1


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(H, 2), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, xy, bf16_ops=True, keep=False):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    xy = np.asarray(xy, np.float64).reshape(-1, 2)
    if bf16_ops:
        # the device computes z1 in float32 as (w0*x + w1*y) + b without contraction
        f32 = np.float32
        x32 = xy.astype(f32)
        pw = np.asarray(p, f32)
        z1 = ((pw[W1:B1:2][None, :] * x32[:, :1] + pw[W1 + 1:B1:2][None, :] * x32[:, 1:2])
              + pw[B1:W2][None, :]).astype(np.float64)
        h1 = bf16(z1 * _sig(z1))
        w2u = bf16(np.asarray(p, np.float32)[W2:B2]).reshape(H, H)
    else:
        z1 = xy @ w1.T + b1
        h1 = z1 * _sig(z1)
        w2u = w2
    z2 = h1 @ w2u.T + b2
    h2 = z2 * _sig(z2)
    out = h2 @ w3 + b3
    if keep:
        return out, dict(xy=xy, z1=z1, h1=h1, w2u=w2u, z2=z2, h2=h2)
    return out


def gradient(p, xy, coef, bf16_ops=True):
    """sum_i coef_i * d f(x_i) / d p, in the flat parameter layout."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    out, c = forward(p, xy, bf16_ops, keep=True)
    coef = np.asarray(coef, np.float64)
    z1, h1, z2, h2, xy = c["z1"], c["h1"], c["z2"], c["h2"], c["xy"]
    s2 = _sig(z2)
    d2 = coef[:, None] * w3[None, :] * (s2 * (1 + z2 * (1 - s2)))
    d2u = bf16(d2) if bf16_ops else d2
    g = np.zeros(NPARAM)
    g[W3:B3] = coef @ h2
    g[B2:W3] = d2.sum(0)
    g[B3] = coef.sum()
    g[W2:B2] = (d2u.T @ h1).ravel()
    dh1 = d2u @ c["w2u"]
    s1 = _sig(z1)
    dz1 = dh1 * (s1 * (1 + z1 * (1 - s1)))
    g[W1:B1] = (dz1.T @ xy).ravel()
    g[B1:W2] = dz1.sum(0)
    return g


def mle_loss(p, xy_data, xy_unif, bf16_ops=True):
    """-mean_data f + log mean_unif exp f, and the per-sample output coefficients."""
    fd = forward(p, xy_data, bf16_ops)
    fu = forward(p, xy_unif, bf16_ops)
    m = fu.max()
    e = np.exp(fu - m)
    loss = -fd.mean() + m + np.log(e.mean())
    coef = np.concatenate([np.full(fd.size, -1.0 / fd.size), e / e.sum()])
    return loss, coef
