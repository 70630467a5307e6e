"""GPU checks of the learned field step (include/ipic_nop.h) through the C ABI: forward,
loss, gradients and Adam against the torch CPU reference (oracle/nop_ref.py, float64 with
the HIP path's BF16 rounding points), the operator inside the step (fallback and audit),
and a trained operator against the method without it."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import nop_ref as R

pytestmark = pytest.mark.gpu

N = (8, 8, 16)
V = N[0] * N[1] * N[2]
SCALES = {"in_E": 1.0, "in_B": 1.0, "in_J": 1.0, "out_E": 0.7, "out_P": 0.3}


def _fo():
    from allscale_ipic3d_amd import field_operator as FO
    return FO


def _params(seed):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed, out_gain=1.0, res_gain=1.0)
    for wo, bo, ci, co in R.OFFSETS:
        p[bo:bo + co] = rng.normal(0, 0.1, size=co).astype(np.float32)
    wo, bo, ci, co = R.OFFSETS[-1]
    p[bo + 6:bo + co] = 0.0
    return p


def _u(seed):
    rng = np.random.default_rng(seed + 100)
    u = np.zeros((V, 16), dtype=np.float32)
    u[:, :9] = rng.normal(size=(V, 9))
    return R.bf(torch.as_tensor(u)).numpy()


def _upload_u(FO, u):
    bits = (u.view(np.uint32) >> 16).astype(np.uint16)   # exact: u holds BF16 values
    d = FO.Dev(bits.nbytes)
    d.upload(bits)
    return d


def _f4(FO, a):
    b = np.zeros((a.shape[0], 4), dtype=np.float32)
    b[:, :3] = a
    d = FO.Dev(b.nbytes)
    d.upload(b)
    return d


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_forward_matches_reference():
    FO = _fo()
    p, u = _params(1), _u(1)
    net = FO.Network(N, SCALES, params=p, train=True)
    du = _upload_u(FO, u)
    net.forward(du.ptr, train=False)
    o_inf = net.out.download(np.float32).reshape(V, 16)
    net.forward(du.ptr, train=True)
    o_tr = net.out.download(np.float32).reshape(V, 16)
    ref = R.from_grid(R.forward(p, u.astype(np.float64), N)).numpy()
    print("forward rel err", _rel(o_inf, ref))
    np.testing.assert_array_equal(o_inf, o_tr)
    # BF16 activations: a value near a rounding tie may round the other way in FP32 than in
    # FP64 (one BF16 ulp, 2^-8), which spreads through the later layers
    assert _rel(o_inf, ref) < 6e-3
    assert np.all(o_inf[:, 6:] == 0.0)


def test_loss_and_gradients_match_reference():
    FO = _fo()
    p, u = _params(2), _u(2)
    rng = np.random.default_rng(7)
    tE = rng.normal(size=(V, 3)).astype(np.float32)
    tB = rng.normal(size=(V, 3)).astype(np.float32)
    net = FO.Network(N, SCALES, params=p, train=True)
    du, dtE, dtB = _upload_u(FO, u), _f4(FO, tE), _f4(FO, tB)
    loss = net.loss_and_grad(dtE.ptr, dtB.ptr, 1.3, 0.7, u_ptr=du.ptr)
    g = net.grad.download(np.float32)

    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    W = []
    for wo, bo, ci, co in R.OFFSETS:
        W.append((R.bf(pt[wo:bo].reshape(3, 3, 3, co, ci).permute(3, 4, 0, 1, 2)), pt[bo:bo + co]))
    # BF16 rounding of weights is applied in the graph; its gradient is the identity
    o = R.forward(None, u.astype(np.float64), N, bf16=True, bf16_grads=True, weights=W)
    L = R.loss(o, SCALES, tE.astype(np.float64), tB.astype(np.float64), 1.3, 0.7, N, (1.0, 1.0, 1.0))
    L.backward()
    gr = pt.grad.numpy()
    print("loss", loss, L.item())
    assert abs(loss - L.item()) <= 2e-3 * abs(L.item())
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    print("gradient cosine", cos)
    assert cos > 0.999
    for l, (wo, bo, ci, co) in enumerate(R.OFFSETS):
        ew, eb = _rel(g[wo:bo], gr[wo:bo]), _rel(g[bo:bo + co], gr[bo:bo + co])
        print(f"layer {l}: weight grad rel {ew:.2e}, bias grad rel {eb:.2e}")
        assert ew < 3e-2 and eb < 3e-2
    # padded entries stay exactly zero
    for (wo, bo, ci, co), (rci, rco) in zip(R.OFFSETS, R.REAL_IO):
        w = g[wo:bo].reshape(27, co, ci)
        assert np.all(w[:, rco:, :] == 0) and np.all(w[:, :, rci:] == 0)
        assert np.all(g[bo + rco:bo + co] == 0)


def test_adam_matches_reference():
    FO = _fo()
    rng = np.random.default_rng(3)
    p = _params(3)
    net = FO.Network(N, SCALES, params=p, train=True)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    q = p.astype(np.float64)
    for step in range(1, 4):
        g = rng.normal(size=p.size).astype(np.float32)
        net.grad.upload(g)
        net.adam(1e-3)
        q, m, v = R.adam(q, g.astype(np.float64), m, v, step, 1e-3, 0.9, 0.999, 1e-8)
    got = net.get_params()
    np.testing.assert_allclose(got, q, rtol=1e-5, atol=1e-7)


def test_apply_matches_reference():
    FO = _fo()
    p, u = _params(4), _u(4)
    rng = np.random.default_rng(11)
    E = rng.normal(size=(V, 3)).astype(np.float32)
    B = rng.normal(size=(V, 3)).astype(np.float32)
    net = FO.Network(N, SCALES, params=p, train=False)
    du, dE, dB, psi = _upload_u(FO, u), _f4(FO, E), _f4(FO, B), FO.Dev(16 * V)
    net.forward(du.ptr)
    o = net.out.download(np.float32).reshape(V, 16)
    net.apply(dE.ptr, dB.ptr, psi.ptr, cell_width=(1.0, 2.0, 0.5))
    e = dE.download(np.float32).reshape(V, 4)[:, :3]
    b = dB.download(np.float32).reshape(V, 4)[:, :3]
    ot = R.to_grid(o.astype(np.float64), N, 16)
    de, db = R.increments(ot, SCALES, (1.0, 0.5, 2.0))
    np.testing.assert_allclose(e, E + R.from_grid(de.unsqueeze(0)).numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(b, B + R.from_grid(db.unsqueeze(0)).numpy(), rtol=1e-5, atol=1e-5)


def _sim(FO, seed, substeps=2, n=16, ppc=8):
    return FO.config2_simulation(n, ppc, seed, field_substeps=substeps)


def test_operator_fallback_reproduces_leapfrog():
    """A check that always fails sends every field phase to leapfrog: the run equals the run
    without an operator bit for bit."""
    FO = _fo()
    a, b = _sim(FO, 5), _sim(FO, 5)
    zero = np.zeros(FO.param_count(), dtype=np.float32)
    a.set_field_operator(zero, SCALES, energy_tol=-1.0)
    sa = a.step(3)
    b.step(3)
    assert sa["fallback_steps"] == 3 and sa["operator_steps"] == 0
    fa, fb = a.download_fields(), b.download_fields()
    for k in ("E", "Bc", "J"):
        np.testing.assert_array_equal(fa[k], fb[k])
    for s in range(2):
        # the order within a cell follows the arrival atomics: compare each cell's particles as sets
        pa, pb = _by_cell(a.download_particles(s)), _by_cell(b.download_particles(s))
        np.testing.assert_array_equal(pa, pb)
    a.close(); b.close()


def _by_cell(p):
    rows = np.concatenate([p["cell"][:, None].astype(np.float64), p["pos"], p["vel"]], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def test_operator_accepted_and_audited():
    """A zero operator (no field change) with open tolerances is applied every step; audits
    report |F_op - F_lf| / |F_lf - F_before| = 1 for it and, with audit_tol = 0, hand the
    step back to leapfrog."""
    FO = _fo()
    a = _sim(FO, 6)
    zero = np.zeros(FO.param_count(), dtype=np.float32)
    a.set_field_operator(zero, SCALES, divb_tol=1e30, energy_tol=1e30)
    f0 = a.download_fields()
    st = a.field_step(use_operator=True)
    f1 = a.download_fields()
    assert st.operator_steps == 1 and st.fallback_steps == 0
    np.testing.assert_array_equal(f0["E"], f1["E"])
    np.testing.assert_array_equal(f0["Bc"], f1["Bc"])
    assert st.max_divb_residual == 0.0
    a.particle_step()
    a.set_field_operator(zero, SCALES, divb_tol=1e30, energy_tol=1e30, audit_every=1, audit_tol=0.0)
    st = a.field_step(use_operator=True)
    assert st.audits == 1 and st.fallback_steps == 1 and st.operator_steps == 0
    assert abs(st.max_audit_error_increment - 1.0) < 1e-6
    a.close()


@pytest.mark.slow
def test_trained_operator_beats_persistence():
    """Trained for a few hundred Adam steps on pairs from two runs, the operator predicts the
    leapfrog increments of a held-out run far better than leaving the field unchanged."""
    FO = _fo()
    n = (16, 16, 16)
    snaps = FO.generate(n, seeds=[1, 2, 3], pairs_per_seed=16, ppc=16, field_substeps=8)
    scales, varE, varB = FO.fit_scales(snaps, dt_field=0.05)
    net = FO.Network(n, scales, params=FO.init_params(0), train=True)
    u_all = FO.pack_all(net, snaps)
    test_idx = list(range(32, 48))
    before = FO.evaluate(net, snaps, u_all, test_idx)
    losses = FO.train(net, snaps, u_all, list(range(32)), varE, varB, epochs=12, lr=2e-3)
    after = FO.evaluate(net, snaps, u_all, test_idx)
    print("scales", scales, "untrained", before, "trained", after, "loss", losses[0], losses[-1])
    assert np.isfinite(losses[-1]) and losses[-1] < 0.25 * losses[0]
    assert after["rel_increment"] < 0.5
    assert after["rel_total"] < after["persistence_rel_total"]
