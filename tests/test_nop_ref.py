"""Pins the test-only reference of the learned field step (oracle/nop_ref.py) without a GPU:
the tap convention by brute force, its gradients by finite differences, the head's curl
against the oracle's leapfrog B pass (field.h:366-371), the curl adjoint used by the loss
gradient, and Adam's first step against its closed form."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import nop_ref as R


def test_param_layout_size():
    assert R.NPARAM == sum(27 * ci * co + co for ci, co in R.LAYERS)


@pytest.mark.parametrize("tap", [0, 4, 13, 17, 22, 26])
def test_conv_tap_convention(tap):
    """W[tap] multiplies x[v + (dx, dy, dz)] with tap = (dx+1)*9 + (dy+1)*3 + (dz+1)."""
    rng = np.random.default_rng(tap)
    n = (4, 5, 6)
    x = torch.as_tensor(rng.normal(size=(1, 1) + n))
    p = np.zeros(R.NPARAM, dtype=np.float32)
    wo, bo, ci, co = R.OFFSETS[0]
    w = np.zeros((27, co, ci), dtype=np.float32)
    w[tap, 0, 0] = 1.0
    p[wo:bo] = w.reshape(-1)
    W, _ = R.unpack(p, bf16=False)[0]
    y = R.conv(torch.cat([x] + [torch.zeros_like(x)] * 15, dim=1), W, None)[0, 0].numpy()
    dx, dy, dz = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
    want = np.zeros(n)
    for a in range(n[0]):
        for b in range(n[1]):
            for c in range(n[2]):
                want[a, b, c] = x[0, 0, (a + dx) % n[0], (b + dy) % n[1], (c + dz) % n[2]]
    np.testing.assert_array_equal(y, want)


def _case(n, seed):
    rng = np.random.default_rng(seed)
    V = n[0] * n[1] * n[2]
    p = R.init_params(seed, out_gain=1.0, res_gain=1.0).astype(np.float64)
    for wo, bo, ci, co in R.OFFSETS:
        p[bo:bo + co] = rng.normal(0, 0.1, size=co)
    # keep padded channels of the biases at zero as the HIP layout does
    wo, bo, ci, co = R.OFFSETS[-1]
    p[bo + 6:bo + co] = 0.0
    u = np.zeros((V, 16))
    u[:, :9] = rng.normal(size=(V, 9))
    tE = rng.normal(size=(V, 3))
    tB = rng.normal(size=(V, 3))
    scales = {"in_E": 1.0, "in_B": 1.0, "in_J": 1.0, "out_E": 0.7, "out_P": 0.3}
    return p, u, tE, tB, scales


def test_reference_gradient_matches_finite_differences():
    n = (4, 4, 4)
    p, u, tE, tB, scales = _case(n, 3)
    iw = (1.0, 0.5, 2.0)

    def L(pp):
        return R.loss(R.forward(pp, u, n, bf16=False), scales, tE, tB, 1.3, 0.7, n, iw)

    pt = torch.tensor(p, requires_grad=True)
    # unpack from a differentiable vector
    def Lt(pt_):
        W = []
        for wo, bo, ci, co in R.OFFSETS:
            W.append((pt_[wo:bo].reshape(3, 3, 3, co, ci).permute(3, 4, 0, 1, 2), pt_[bo:bo + co]))
        o = R.forward(None, u, n, bf16=False, weights=W)
        return R.loss(o, scales, tE, tB, 1.3, 0.7, n, iw)

    Lt(pt).backward()
    g = pt.grad.numpy()
    rng = np.random.default_rng(0)
    # real (non-padded) entries of several layers, and biases
    picks = []
    for l, ((wo, bo, ci, co), (rci, rco)) in enumerate(zip(R.OFFSETS, R.REAL_IO)):
        for _ in range(2):
            tap, a, b = rng.integers(27), rng.integers(rco), rng.integers(rci)
            picks.append(wo + (tap * co + a) * ci + b)
        picks.append(bo + rng.integers(rco))
    for i in picks:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (L(pp).item() - L(pm).item()) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_curl_adjoint():
    rng = np.random.default_rng(1)
    n = (5, 6, 7)
    iw = (1.0, 0.5, 0.25)
    psi = torch.as_tensor(rng.normal(size=(3,) + n))
    g = torch.as_tensor(rng.normal(size=(3,) + n))
    lhs = (R.curl_plus(psi, iw) * g).sum().item()
    rhs = (psi * R.curl_minus(g, iw)).sum().item()
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_head_curl_is_the_leapfrog_b_pass():
    """With psi = dt E, -curl+ psi is the Bc change of one leapfrog substep from Bc = 0, J = 0
    (the oracle's two-pass E then B update leaves E unchanged then)."""
    n = (4, 5, 6)
    w = (1.0, 1.5, 2.0)
    dt = 0.1
    rng = np.random.default_rng(2)
    E = rng.normal(size=(n[0] * n[1] * n[2], 3))
    u = O.Universe(n, w=w, dt=dt)
    u.set_fields(E, np.zeros_like(E))
    u.set_J(np.zeros_like(E))
    u.field_solve()
    f = u.fields()
    u.close()
    psi = R.to_grid(dt * E, n, 3)[0]
    dB = -R.curl_plus(psi, [1.0 / v for v in w])
    got = R.from_grid(dB.unsqueeze(0)).numpy()
    np.testing.assert_allclose(got, f["Bc"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f["E"], E, rtol=0, atol=0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    q, m, v = R.adam(p, g, np.zeros(100), np.zeros(100), 1, 1e-2, 0.9, 0.999, 1e-8)
    np.testing.assert_allclose(q, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12)
