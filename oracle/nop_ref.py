"""Test-only reference of the learned field step (include/ipic_nop.h), for checking the HIP
kernels: the same network, head and loss written with torch CPU conv3d (circular padding) in
float64, with BF16 rounding applied where the HIP path rounds (input, weights, the stored
activations and, optionally, the gradients fed to the data/weight-gradient convolutions).

Only tests and the benchmark's checks import this module.  The reference (allscale_ipic3d)
has no learned field step; the physics it must agree with is the leapfrog B pass of
field.h:366-371, restated in the C oracle (or_field_solve) and used to pin the head below.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [(16, 32)] + [(32, 32)] * 8 + [(32, 16)]   # padded (cin, cout) per layer


def layer_offsets():
    offs, o = [], 0
    for ci, co in LAYERS:
        offs.append((o, o + 27 * ci * co, ci, co))
        o += 27 * ci * co + co
    return offs, o


OFFSETS, NPARAM = layer_offsets()
REAL_IO = [(9, 32)] + [(32, 32)] * 8 + [(32, 6)]    # channels that carry data


def init_params(seed=0, out_gain=0.1, res_gain=0.1):
    """He-normal weights over the real channels (padding stays zero), zero biases; the last
    conv of each residual block and the projection start small."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    for l, ((wo, bo, ci, co), (rci, rco)) in enumerate(zip(OFFSETS, REAL_IO)):
        std = np.sqrt(2.0 / (27 * rci))
        if l == len(LAYERS) - 1:
            std *= out_gain
        elif l >= 1 and l % 2 == 0:
            std *= res_gain
        w = np.zeros((27, co, ci), dtype=np.float32)
        w[:, :rco, :rci] = rng.normal(0.0, std, size=(27, rco, rci))
        p[wo:bo] = w.reshape(-1)
    return p


def bf(x):
    return x.to(torch.bfloat16).to(x.dtype)


def unpack(params, dtype=torch.float64, bf16=True):
    """[(W[co][ci][3][3][3], b[co])] from the flat vector (tap = (dx+1)*9 + (dy+1)*3 + (dz+1))."""
    p = torch.as_tensor(np.asarray(params)).to(dtype)
    out = []
    for wo, bo, ci, co in OFFSETS:
        w = p[wo:bo].reshape(3, 3, 3, co, ci).permute(3, 4, 0, 1, 2).contiguous()
        if bf16:
            w = bf(w)
        out.append((w, p[bo:bo + co]))
    return out


def conv(x, w, b):
    """3x3x3 periodic cross-correlation: y[v] = sum_tap W[tap] x[v + offset(tap)] + b."""
    return F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), w, b)


def to_grid(a, n, c):
    """[V][c] (row-major lattice) -> [1][c][nx][ny][nz]"""
    t = torch.as_tensor(np.asarray(a)).reshape(n[0], n[1], n[2], c)
    return t.permute(3, 0, 1, 2).unsqueeze(0)


def from_grid(t):
    c = t.shape[1]
    return t[0].permute(1, 2, 3, 0).reshape(-1, c)


class _BfGrad(torch.autograd.Function):
    """Identity whose backward rounds the incoming gradient to BF16 (the HIP backward feeds
    BF16 gradients to its convolutions)."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return bf(g)


def forward(params, u, n, dtype=torch.float64, bf16=True, bf16_grads=False, weights=None):
    """u: [V][16] (already BF16 values when bf16); returns o as [1][16][nx][ny][nz].
    bf16_grads rounds the gradient arriving at every convolution's output, as the HIP backward
    does before its data- and weight-gradient convolutions."""
    W = weights if weights is not None else unpack(params, dtype, bf16)
    r = (lambda t: bf(t)) if bf16 else (lambda t: t)
    g_ = (lambda t: _BfGrad.apply(t)) if bf16_grads else (lambda t: t)
    x = g_(conv(to_grid(u, n, 16).to(dtype), *W[0]))
    for b in range(4):
        a = r(F.silu(x))
        h = g_(conv(a, *W[1 + 2 * b]))
        g = r(F.silu(h))
        x = x + g_(conv(g, *W[2 + 2 * b]))
    a = r(F.silu(x))
    return g_(conv(a, *W[9]))


def pack(dE, dBc, J, scales, bf16=True):
    """[V][16] input of the operator from [V][3] fields."""
    V = dE.shape[0]
    u = np.zeros((V, 16), dtype=np.float64)
    u[:, 0:3] = dE / scales["in_E"]
    u[:, 3:6] = dBc / scales["in_B"]
    u[:, 6:9] = J / scales["in_J"]
    t = torch.as_tensor(u)
    return (bf(t.to(torch.float32)).to(torch.float64) if bf16 else t).numpy()


def curl_plus(psi, iw):
    """Forward-difference curl (the B pass of field.h:366-371), psi: [3][nx][ny][nz]."""
    px, py, pz = psi[0], psi[1], psi[2]
    d = lambda f, ax: (torch.roll(f, -1, dims=ax) - f) * iw[ax]
    return torch.stack([d(pz, 1) - d(py, 2), d(px, 2) - d(pz, 0), d(py, 0) - d(px, 1)])


def curl_minus(g, iw):
    """Backward-difference curl (the E pass of field.h:359-363), the adjoint of curl_plus."""
    gx, gy, gz = g[0], g[1], g[2]
    d = lambda f, ax: (f - torch.roll(f, 1, dims=ax)) * iw[ax]
    return torch.stack([d(gz, 1) - d(gy, 2), d(gx, 2) - d(gz, 0), d(gy, 0) - d(gx, 1)])


def increments(o, scales, iw):
    """(dE_pred, dB_pred) as [3][nx][ny][nz] each from o ([1][16][...])."""
    dE = scales["out_E"] * o[0, 0:3]
    psi = scales["out_P"] * o[0, 3:6]
    return dE, -curl_plus(psi, iw)


def loss(o, scales, tE, tB, varE, varB, n, iw):
    """Mean over the lattice of |dE - tE|^2 / varE + |dB - tB|^2 / varB (tE, tB: [V][3])."""
    dE, dB = increments(o, scales, iw)
    te = to_grid(tE, n, 3)[0].to(o.dtype)
    tb = to_grid(tB, n, 3)[0].to(o.dtype)
    V = n[0] * n[1] * n[2]
    return ((dE - te) ** 2).sum() / (varE * V) + ((dB - tb) ** 2).sum() / (varB * V)


def adam(p, g, m, v, step, lr, b1, b2, eps):
    """The Adam update of ipic_nop_adam (bias-corrected), in float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    return p - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v
