"""Pins the FP64 oracle against the reference's own known-answer tests, its golden
trajectory files and closed forms.  CPU only.

Reference citations: code/app/test/*.cc and code/app/include/ipic3d/app/*.h.
"""
import math
import os

import numpy as np
import pytest

import oracle as O


def _load_rows(golden_dir, name):
    return np.loadtxt(os.path.join(golden_dir, name))


# ---------------------------------------------------------------- particle.h ------------

def test_position_update_exact():
    # particle_test.cc:7-42
    p = O.make_particles([[0.0, 0.0, 0.0]], [[1.0, 2.0, 3.0]], 1.0, 1.0)
    O.lib().or_update_position(O._p(p), 1.0)
    assert list(p["pos"][0]) == [1.0, 2.0, 3.0]
    O.lib().or_update_position(O._p(p), 0.5)
    assert list(p["pos"][0]) == [1.5, 3.0, 4.5]


def test_boris_kat():
    # particle_test.cc:55-79: no force -> unchanged; E=(1,-1,3), B=(0,1,2), qom=4, dt=1
    v = O.boris([1.0, 2.0, 3.0], [0, 0, 0], [0, 0, 0], 4.0, 1.0)
    assert list(v) == [1.0, 2.0, 3.0]
    v = O.boris([1.0, 2.0, 3.0], [1.0, -1.0, 3.0], [0.0, 1.0, 2.0], 4.0, 1.0)
    assert v == pytest.approx([-2.42857, 3.71429, 12.14286], abs=1e-4)
    # closed form of the same numbers: -17/7, 26/7, 85/7
    assert v == pytest.approx([-17 / 7, 26 / 7, 85 / 7], rel=1e-14)


def test_boris_uniform_B_closed_form():
    # E=0, uniform B: |v| is conserved and the velocity rotates by 2*atan(qom*B*dt/2) per step
    rng = np.random.default_rng(3)
    for _ in range(20):
        v0 = rng.normal(size=3)
        qom, B, dt = rng.uniform(-3, 3), rng.uniform(0.1, 2.0), rng.uniform(0.01, 0.5)
        v = O.boris(v0, [0, 0, 0], [0, 0, B], qom, dt)
        assert np.linalg.norm(v) == pytest.approx(np.linalg.norm(v0), rel=1e-13)
        ang0 = math.atan2(v0[1], v0[0])
        ang1 = math.atan2(v[1], v[0])
        expect = -2.0 * math.atan(qom * B * dt / 2.0)
        d = (ang1 - ang0 - expect + math.pi) % (2 * math.pi) - math.pi
        assert abs(d) < 1e-12
        assert v[2] == pytest.approx(v0[2], abs=1e-15)


def _run_golden(E, B, q, qom, pos0, vel0, dt, nsteps):
    p = O.make_particles([pos0], [vel0], q, qom)
    Ep, Bp = O.f64(E), O.f64(B)
    # push velocity back by half a step (simulation_test.cc:236, 376)
    O.lib().or_update_velocity(O._p(p), O._p(Ep), O._p(Bp), -0.5 * dt)
    traj = []
    for _ in range(nsteps):
        O.lib().or_update_velocity(O._p(p), O._p(Ep), O._p(Bp), dt)
        O.lib().or_update_position(O._p(p), dt)
        traj.append(np.concatenate([p["pos"][0], p["vel"][0]]))
    return np.array(traj)


def test_larmor_golden(golden_dir):
    # simulation_test.cc:191-262 and results/larmor_radius.txt (row k = after k+1 steps)
    q, mass = -1.602e-19, 9.109e-31
    rL = mass * 1e5 / (abs(q) * 0.01)
    assert rL == pytest.approx(5.686e-05, abs=1e-6)
    traj = _run_golden([0, 0, 0], [0, 0, 0.01], q, q / mass, [rL, 0, 0], [0, 1e5, 0], 3e-11, 1000)
    rows = _load_rows(golden_dir, "larmor_radius_rows.txt")
    for r in rows:
        k = int(r[0])
        got = traj[k]
        # the file carries 6 significant digits
        np.testing.assert_allclose(got[[0, 1, 3, 4]], r[[2, 3, 5, 6]], rtol=2e-5, atol=0)
    # the asserts of simulation_test.cc:255-261
    last = traj[-1]
    assert last[0] == pytest.approx(-4.50134e-05, abs=1e-10)
    assert last[1] == pytest.approx(3.47983e-05, abs=1e-9)
    assert last[3] == pytest.approx(-63242.7, abs=10)
    assert last[4] == pytest.approx(-77462.0, abs=10)


def test_exb_golden(golden_dir):
    # simulation_test.cc:340-407 and results/exb_drift.txt
    traj = _run_golden([0.2, 0, 0], [0, 0, 1.0], 1.0, 1.0, [0, 0, 0], [0.5, 0, 0], 0.01, 5000)
    rows = _load_rows(golden_dir, "exb_drift_rows.txt")
    for r in rows:
        k = int(r[0])
        got = traj[k]
        np.testing.assert_allclose(got[[0, 1, 3, 4]], r[[2, 3, 5, 6]], rtol=3e-5, atol=3e-6)
    # the E x B drift velocity E x B / |B|^2 = (0, -0.2, 0): mean vy over whole gyrations
    assert traj[:, 4].mean() == pytest.approx(-0.2, abs=5e-3)


# ---------------------------------------------------------------- cell.h gather ---------

def test_trilinear_kats():
    # field_test.cc:231-273
    ones = np.ones(8)
    assert O.lib().or_trilinear(O._p(ones), O._p(O.f64([0.5, 0.5, 0.5])), 1.0) == 1.0
    lin = O.f64([i + 2 * j + 3 * k for i in range(2) for j in range(2) for k in range(2)])
    assert O.lib().or_trilinear(O._p(lin), O._p(O.f64([0.5, 0.5, 0.5])), 0.75) == 4.0
    assert O.lib().or_trilinear(O._p(lin), O._p(O.f64([0.0, 0.0, 0.0])), 3.0) == 0.0
    assert O.lib().or_trilinear(O._p(lin), O._p(O.f64([1.0, 1.0, 1.0])), 1.0) == 6.0


def test_trilinear_exact_on_trilinear_functions():
    rng = np.random.default_rng(0)
    for _ in range(50):
        a = rng.normal(size=8)  # f = sum a_m * monomial_m(x,y,z) over {1,x,y,z,xy,xz,yz,xyz}

        def f(x, y, z):
            return (a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * y + a[5] * x * z + a[6] * y * z
                    + a[7] * x * y * z)
        corners = O.f64([f(i, j, k) for i in range(2) for j in range(2) for k in range(2)])
        pos = rng.uniform(size=3)
        got = O.lib().or_trilinear(O._p(corners), O._p(O.f64(pos)), 1.0)
        assert got == pytest.approx(f(*pos), abs=1e-12)


# ---------------------------------------------------------------- field.h stencils ------

def _grid(dims):
    return np.zeros(tuple(dims) + (3,), dtype=np.float64)


def test_curlB_kat():
    # field_test.cc:45-90
    dims = O.i64([2, 2, 2])
    Bc = _grid(dims)
    pos = O.i64([1, 1, 1])
    w = O.f64([1, 1, 1])
    curl = np.zeros(3)
    O.lib().or_ref_curlB(O._p(dims), O._p(Bc), O._p(pos), O._p(w), O._p(curl))
    assert list(curl) == [0, 0, 0]
    for i in range(2):
        for j in range(2):
            for k in range(2):
                Bc[1 - i, 1 - j, 1 - k] = [0.0, 2.0, i + 2.0 * j + 3.0 * k]
    O.lib().or_ref_curlB(O._p(dims), O._p(Bc), O._p(pos), O._p(w), O._p(curl))
    assert curl == pytest.approx([-2.0, 1.0, 0.0], abs=1e-6)
    w = O.f64([1.0, 5.0, 10.0])
    O.lib().or_ref_curlB(O._p(dims), O._p(Bc), O._p(pos), O._p(w), O._p(curl))
    assert curl == pytest.approx([-0.4, 1.0, 0.0], abs=1e-6)


def test_curlE_kat():
    # field_test.cc:93-142
    dims = O.i64([2, 2, 2])
    E = _grid(dims)
    pos = O.i64([0, 0, 0])
    w = O.f64([1, 1, 1])
    curl = np.zeros(3)
    O.lib().or_ref_curlE(O._p(dims), O._p(E), O._p(pos), O._p(w), O._p(curl))
    assert list(curl) == [0, 0, 0]
    for i in range(2):
        for j in range(2):
            for k in range(2):
                E[i, j, k] = [0.0, 2.0, i + 2.0 * j + 3.0 * k]
    O.lib().or_ref_curlE(O._p(dims), O._p(E), O._p(pos), O._p(w), O._p(curl))
    assert curl == pytest.approx([2.0, -1.0, 0.0], abs=1e-6)
    w = O.f64([1.0, 5.0, 10.0])
    O.lib().or_ref_curlE(O._p(dims), O._p(E), O._p(pos), O._p(w), O._p(curl))
    assert curl == pytest.approx([0.4, -1.0, 0.0], abs=1e-6)


def test_interp_kats():
    # field_test.cc:145-229
    bdims, ndims = O.i64([2, 2, 2]), O.i64([3, 3, 3])
    Bc = _grid(bdims) + 1.0
    B = _grid(ndims)
    pos = O.i64([1, 1, 1])
    O.lib().or_ref_interpC2N(O._p(bdims), O._p(Bc), O._p(ndims), O._p(B), O._p(pos))
    assert list(B[1, 1, 1]) == pytest.approx([1, 1, 1])
    for i in range(2):
        for j in range(2):
            for k in range(2):
                Bc[1 - i, 1 - j, 1 - k] = [0.0, (i + 1.0) * (j + 1.0) * (k + 1.0), i + 2.0 * j + 3.0 * k]
    O.lib().or_ref_interpC2N(O._p(bdims), O._p(Bc), O._p(ndims), O._p(B), O._p(pos))
    assert list(B[1, 1, 1]) == pytest.approx([0.0, 3.375, 3.0], abs=1e-6)

    ndims, bdims = O.i64([2, 2, 2]), O.i64([1, 1, 1])
    B = _grid(ndims) + 1.0
    Bc = _grid(bdims)
    pos = O.i64([0, 0, 0])
    O.lib().or_ref_interpN2C(O._p(ndims), O._p(B), O._p(bdims), O._p(Bc), O._p(pos))
    assert list(Bc[0, 0, 0]) == pytest.approx([1, 1, 1])
    for i in range(2):
        for j in range(2):
            for k in range(2):
                B[i, j, k] = [0.0, 2.0, i + 2.0 * j + 3.0 * k]
    O.lib().or_ref_interpN2C(O._p(ndims), O._p(B), O._p(bdims), O._p(Bc), O._p(pos))
    assert list(Bc[0, 0, 0]) == pytest.approx([0.0, 2.0, 3.0], abs=1e-6)


def test_update_fields_on_boundaries_kat():
    # field_test.cc:316-416: size 10 -> Field 13^3, BcField 12^3
    fs, bs = 13, 12
    idx = np.stack(np.meshgrid(np.arange(fs), np.arange(fs), np.arange(fs), indexing="ij"), -1).astype(float)
    E = idx.copy()
    B = idx.copy()
    bidx = np.stack(np.meshgrid(np.arange(bs), np.arange(bs), np.arange(bs), indexing="ij"), -1).astype(float)
    Bc = bidx.copy()
    E[1:-1, 1:-1, 1:-1] = 1.0
    B[1:-1, 1:-1, 1:-1] = 2.0
    Bc[1:-1, 1:-1, 1:-1] = 3.0
    assert np.all(E.sum(axis=(0, 1, 2)) == 6527.0)
    assert np.all(B.sum(axis=(0, 1, 2)) == 7858.0)
    assert np.all(Bc.sum(axis=(0, 1, 2)) == 7004.0)
    E, B, Bc = (np.ascontiguousarray(a) for a in (E, B, Bc))
    O.lib().or_ref_update_boundaries(fs, O._p(E), O._p(B), bs, O._p(Bc))
    fS, bS = fs - 1, bs - 1
    for i in range(2):
        for j in range(2):
            for k in range(2):
                fp = (i * fS, j * fS, k * fS)
                bp = (i * bS, j * bS, k * bS)
                assert list(E[fp]) == list(map(float, fp))
                assert list(B[fp]) == list(map(float, fp))
                assert list(Bc[bp]) == list(map(float, bp))
    for i in range(1, fS):
        for j in range(1, fS):
            for p in [(0, i, j), (fS, i, j), (i, 0, j), (i, fS, j), (i, j, 0), (i, j, fS)]:
                assert list(E[p]) == [1.0, 1.0, 1.0]
    E[1:-1, 1:-1, 1:-1] = idx[1:-1, 1:-1, 1:-1]
    B[1:-1, 1:-1, 1:-1] = 2.0 * idx[1:-1, 1:-1, 1:-1]
    O.lib().or_ref_update_boundaries(fs, O._p(E), O._p(B), bs, O._p(Bc))
    assert list(E[0, 5, 11]) == [11.0, 5.0, 11.0]
    assert list(B[9, 12, 2]) == [18.0, 2.0, 4.0]


def test_leapfrog_zero_kat():
    # field_test.cc:515-557: all-zero input stays zero
    fd, bd, jd = O.i64([6, 6, 6]), O.i64([7, 7, 7]), O.i64([6, 6, 6])
    E, Bc, J = _grid(fd), _grid(bd), _grid(jd)
    O.lib().or_ref_leapfrog_node(O._p(fd), O._p(E), O._p(bd), O._p(Bc), O._p(jd), O._p(J),
                                 O._p(O.i64([2, 2, 2])), O._p(O.f64([1, 1, 1])), 1.0, 1.0)
    assert not E.any() and not Bc.any()


def test_dipole_kats():
    # field_test.cc:25-31 with inputs/test.inp: 8 cells over L=10 (dspace 1.25), centre (6,5,5),
    # planet radius (L_square) 0.5, B0=(0,0,1e-4), B1z=2, drift u0=0.02; origin = centre - L/2;
    # node (2,2,2) lies at origin + (1,1,1)*1.25 (initFields, field.h:97-104)
    origin = np.array([6.0, 5.0, 5.0]) - 5.0
    loc = origin + 1.25
    E, Bext, B = np.zeros(3), np.zeros(3), np.zeros(3)
    O.lib().or_dipole_field(O._p(O.f64(loc)), O._p(O.f64([0.02, 0, 0])), O._p(O.f64([0, 0, 1e-4])),
                            O._p(O.f64([6, 5, 5])), 0.5, 2.0, O._p(E), O._p(Bext), O._p(B))
    assert Bext[0] == pytest.approx(-0.0009123559809416308, abs=1e-10)
    assert Bext[1] == pytest.approx(-0.0009123559809416308, abs=1e-10)
    assert Bext[2] == pytest.approx(0.0, abs=1e-15)
    assert B[2] == pytest.approx(0.0001, abs=1e-10)
    # planet radius 1000 (field_test.cc:33-41): every node is inside -> Bext = 0
    O.lib().or_dipole_field(O._p(O.f64(loc)), O._p(O.f64([0.02, 0, 0])), O._p(O.f64([0, 0, 1e-4])),
                            O._p(O.f64([6, 5, 5])), 1000.0, 2.0, O._p(E), O._p(Bext), O._p(B))
    assert not Bext.any()
    # field_test.cc:277-312 with inputs/tiny.inp: 16 cells (dspace 0.625), node (1,2,4)
    loc = origin + np.array([0, 1, 3]) * 0.625
    O.lib().or_dipole_field(O._p(O.f64(loc)), O._p(O.f64([0.02, 0, 0])), O._p(O.f64([0, 0, 1e-4])),
                            O._p(O.f64([6, 5, 5])), 0.5, 2.0, O._p(E), O._p(Bext), O._p(B))
    assert Bext == pytest.approx([-0.00054926697, -0.00048060860, 0.00028836516], abs=1e-10)
    assert B[2] == pytest.approx(0.00038836516, abs=1e-10)
    assert E == pytest.approx([0.0, 2.0e-06, 0.0], abs=1e-8)


# ---------------------------------------------------------------- deposit ---------------

def _kat_particles():
    # cell_test.cc:232-262 / simulation_test.cc:713-737: 8 particles in the 2x2x2 sub-box
    P = {}
    P[0] = ([0.4, 0.4, 0.4], [0.6, 0.6, 0.6])
    P[1] = ([0.8, 0.4, 0.4], [0.3, 0.3, 0.3])
    P[2] = ([0.4, 0.8, 0.4], [0.1, 0.1, 0.1])
    P[3] = ([0.75, 0.75, 0.4], [0.8, 0.8, 0.8])
    P[4] = ([0.4, 0.4, 0.6], [0.6, 0.6, 0.6])
    P[5] = ([0.8, 0.4, 0.8], [0.3, 0.3, 0.3])
    P[6] = ([0.25, 0.75, 0.75], [0.8, 0.8, 0.8])
    P[7] = ([0.75, 0.75, 0.8], [0.8, 0.8, 0.8])
    pos = np.array([P[i][0] for i in range(8)])
    vel = np.array([P[i][1] for i in range(8)])
    return O.make_particles(pos, vel, 1.0, 1.0)


def test_deposit_kat():
    # cell_test.cc:212-291: 2x2x2 cells of width 0.5, J(1,1,1) = 1.0952 per component
    u = O.Universe(2, w=(0.5, 0.5, 0.5), dt=0.1)
    u.add_particles(_kat_particles())
    assert list(u.cell_counts()) == [1] * 8
    u.deposit(nthreads=1)
    J = u.fields()["J"].reshape(2, 2, 2, 3)
    assert J[1, 1, 1] == pytest.approx([1.0952] * 3, abs=1e-4)


def test_deposit_matches_scatter_formulation():
    # Node-centric gather (the reference's form) equals the particle-centric scatter of
    # q v w to the 8 cell corners, divided by vol*8; and sum(J)*vol*8 = sum(q v).
    rng = np.random.default_rng(7)
    n = (3, 4, 5)
    w = (0.5, 0.7, 1.3)
    u = O.Universe(n, w=w, origin=(-1.0, 2.0, 0.5))
    N = 400
    cells = np.stack([rng.integers(0, n[d], N) for d in range(3)], -1)
    off = rng.uniform(size=(N, 3))
    pos = np.array([-1.0, 2.0, 0.5]) + (cells + off) * np.array(w)
    vel = rng.normal(size=(N, 3))
    q = rng.uniform(0.5, 2.0, N)
    parts = O.make_particles(pos, vel, 1.0, -1.0)
    parts["q"] = q
    u.add_particles(parts)
    u.deposit()
    J = u.fields()["J"].reshape(n + (3,))
    ref = np.zeros(n + (3,))
    for p in range(N):
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    fac = ((off[p, 0] if i else 1 - off[p, 0]) * (off[p, 1] if j else 1 - off[p, 1])
                           * (off[p, 2] if k else 1 - off[p, 2]))
                    node = ((cells[p, 0] + i) % n[0], (cells[p, 1] + j) % n[1], (cells[p, 2] + k) % n[2])
                    ref[node] += q[p] * vel[p] * fac
    vol = w[0] * w[1] * w[2]
    ref = ref / vol / 8.0
    np.testing.assert_allclose(J, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(J.sum(axis=(0, 1, 2)) * vol * 8, (q[:, None] * vel).sum(0), rtol=1e-12)


# ---------------------------------------------------------------- leapfrog --------------

def test_leapfrog_divB_invariant_and_energy():
    # div+ of Bc commutes with curl+ (field.h:367-371): it stays constant to roundoff.
    rng = np.random.default_rng(1)
    n = (6, 5, 7)
    u = O.Universe(n, w=(1.0, 0.8, 1.2), dt=0.05, c=1.0)
    E0 = rng.normal(size=(np.prod(n), 3)) * 0.1
    B0 = rng.normal(size=(np.prod(n), 3)) * 0.1
    u.set_fields(E0, B0)
    d0 = u.divB()[1]
    for _ in range(50):
        u.field_solve()
    assert abs(u.divB()[1] - d0) < 1e-12 * max(1.0, d0)


def test_leapfrog_plane_wave_dispersion():
    # In vacuum the two-pass scheme is the Yee scheme on the index lattice: a plane wave
    # E = (0, a cos(k x), 0), Bc = (0, 0, b cos(k (x + 1/2))) oscillates at the discrete
    # frequency sin(w dt/2) = c dt sin(k dx/2)/dx.  Check the E amplitude after n steps.
    N, dx, c, dt = 16, 1.0, 1.0, 0.3
    k = 2 * math.pi * 2 / (N * dx)
    u = O.Universe((N, 1, 1), w=(dx, 1.0, 1.0), dt=dt, c=c)
    x = np.arange(N) * dx
    # choose the eigenmode of the update: E_y at nodes (x), B_z at centres (x + dx/2)
    omega = 2.0 / dt * math.asin(c * dt * math.sin(k * dx / 2) / dx)
    E = np.zeros((N, 3))
    B = np.zeros((N, 3))
    E[:, 1] = np.cos(k * x)
    # B lags E by half a step in time in the leapfrog
    B[:, 2] = -(math.sin(k * dx / 2) / dx) / (math.sin(omega * dt / 2) / dt) * c ** 0 * np.cos(
        k * (x + dx / 2) + omega * dt / 2) * 1.0
    u.set_fields(E, B)
    steps = 40
    for _ in range(steps):
        u.field_solve()
    Ey = u.fields()["E"][:, 1]
    # projection on cos/sin(kx) gives the phase advanced by omega*steps*dt
    a = 2 * np.mean(Ey * np.cos(k * x))
    b = 2 * np.mean(Ey * np.sin(k * x))
    amp = math.hypot(a, b)
    phase = math.atan2(-b, a)
    assert amp == pytest.approx(1.0, abs=1e-10)
    expect = (omega * steps * dt + math.pi) % (2 * math.pi) - math.pi
    assert abs(((phase - expect + math.pi) % (2 * math.pi)) - math.pi) < 1e-9


# ---------------------------------------------------------------- rebin -----------------

def test_migration_one_particle_reflect():
    # simulation_test.cc:410-556: the reference mover sees no field (dipole of moment 0),
    # so the particle streams freely; reflection at the global faces (cell.h:1103-1121).
    u = O.Universe(3, w=(0.5, 0.5, 0.5), dt=0.1, bc="reflect", sink_center=(0, 0, 0), sink_radius=0.0)
    u.add_particles(O.make_particles([[0.46, 0.8, 0.0]], [[1.0, 0.5, 0.5]], 1.0, 1.0))
    u.move(); u.rebin()
    counts = u.cell_counts().reshape(3, 3, 3)
    assert counts[1, 1, 0] == 1 and counts.sum() == 1
    parts, cell = u.particles()
    assert parts["pos"][0][0] == pytest.approx(0.56, abs=1e-3)
    assert parts["vel"][0][0] == pytest.approx(1.0, abs=1e-2)
    for _ in range(100):
        u.move(); u.rebin()
    counts = u.cell_counts().reshape(3, 3, 3)
    assert counts[0, 1, 0] == 1 and counts.sum() == 1
    parts, _ = u.particles()
    assert parts["pos"][0][0] == pytest.approx(0.06, abs=1e-2)
    assert parts["vel"][0][0] == pytest.approx(-1.0, abs=1e-1)


def _placement_ok(u):
    parts, cell = u.particles()
    n = np.array(u.n)
    cc = np.stack(np.unravel_index(cell, u.n), -1)
    w = np.array(list(u.props.w))
    org = np.array(list(u.props.origin))
    rel = parts["pos"] - (org + (cc + 0.5) * w)
    return np.all(np.abs(rel) <= w / 2 + 1e-12)


def test_migration_many_particles_reflect():
    # simulation_test.cc:687-777: 8 particles, 20 steps, count and placement preserved
    u = O.Universe(4, w=(0.5, 0.5, 0.5), dt=0.1, bc="reflect", sink_center=(0, 0, 0), sink_radius=0.0)
    u.add_particles(_kat_particles())
    assert u.num_particles() == 8 and _placement_ok(u)
    for _ in range(20):
        u.move(); u.rebin()
    assert u.num_particles() == 8 and _placement_ok(u)


def test_single_particle_boris_mover_free_stream():
    # simulation_test.cc:135-184: 1 cell, 9 steps of free streaming -> z = 0.9
    u = O.Universe(1, w=(1, 1, 1), dt=0.1, bc="reflect", sink_center=(0, 0, 0), sink_radius=0.0)
    u.add_particles(O.make_particles([[0.5, 0.5, 0.0]], [[0, 0, 1.0]], 1.0, 1.0))
    for _ in range(9):
        u.move(); u.rebin()
    parts, _ = u.particles()
    assert parts["pos"][0] == pytest.approx([0.5, 0.5, 0.899], abs=1e-3)


def test_periodic_rebin_brute_force():
    # Periodic rebin: every particle ends in floor((x-o)/w) mod N, counts are conserved,
    # kept particles precede imports, and the per-cell order follows cell.h:1272-1306.
    rng = np.random.default_rng(11)
    n = (4, 3, 5)
    w = np.array([0.5, 1.0, 0.75])
    org = np.array([1.0, -2.0, 0.0])
    u = O.Universe(n, w=tuple(w), origin=tuple(org), dt=1.0)
    N = 2000
    cells = np.stack([rng.integers(0, n[d], N) for d in range(3)], -1)
    off = rng.uniform(size=(N, 3))
    pos = org + (cells + off) * w
    # displacement of up to +-1.4 cells so some particles move two cells
    disp = rng.uniform(-1.4, 1.4, size=(N, 3)) * w * (rng.uniform(size=(N, 1)) < 0.3)
    vel = disp  # dt = 1: position += velocity
    parts = O.make_particles(pos, vel, 1.0, 1.0)
    parts["q"] = np.arange(N)  # tag
    u.add_particles(parts)
    before, bcell = u.particles()
    # zero fields -> move is a pure drift
    u.move()
    u.rebin()
    after, acell = u.particles()
    assert len(after) == N
    L = np.array(n) * w
    tag = after["q"].astype(int)
    assert sorted(tag) == list(range(N))
    # expected position = wrapped (pos + disp); expected cell = floor
    src = {int(t): i for i, t in enumerate(before["q"].astype(int))}
    for i in range(N):
        t = tag[i]
        b = before[src[t]]
        expect = b["pos"] + b["vel"]
        expect = org + np.mod(expect - org, L)
        np.testing.assert_allclose(after["pos"][i], expect, atol=1e-12)
        ec = np.floor((after["pos"][i] - org) / w).astype(int)
        ec = np.minimum(ec, np.array(n) - 1)
        assert np.ravel_multi_index(tuple(ec), n) == acell[i]
    assert _placement_ok(u)


def test_rebin_order_matches_reference_import_order():
    # cell.h:1272-1306: cell content = kept particles (original order), then arrivals in
    # direction order (0,0,0),(0,0,1),...,(2,2,2) (source = cell + dir - 1), each in the
    # source's order.
    n = (3, 3, 3)
    u = O.Universe(n, dt=1.0)
    # target cell (1,1,1): one kept particle, arrivals from (0,1,1) [dir (0,1,1)] and
    # (2,2,2) [dir (2,2,2)] and (1,1,0) [dir (1,1,0)]
    ps = O.make_particles(
        [[1.5, 1.5, 1.5], [0.9, 1.5, 1.5], [2.1, 2.1, 2.1], [1.5, 1.5, 0.5], [0.8, 1.2, 1.4]],
        [[0, 0, 0], [0.2, 0, 0], [-0.2, -0.2, -0.2], [0, 0, 0.7], [0.3, 0, 0]], 1.0, 1.0)
    ps["q"] = [10, 20, 30, 40, 21]
    u.add_particles(ps)
    u.move(); u.rebin()
    parts, cell = u.particles()
    tags = [int(t) for t, c in zip(parts["q"], cell) if c == 13]
    # kept 10, then dir index order: (0,1,1)-> tags 20,21 (source order), (1,1,0)->40, (2,2,2)->30
    assert tags == [10, 20, 21, 40, 30]


def test_sink_removes_particles():
    u = O.Universe(4, dt=0.1, sink_center=(2.0, 2.0, 2.0), sink_radius=0.5)
    u.add_particles(O.make_particles([[2.3, 2.0, 2.0], [3.5, 3.5, 3.5]], [[-0.5, 0, 0], [0, 0, 0]], 1.0, 1.0))
    u.move()
    assert u.rebin() == 1
    assert u.num_particles() == 1


# ---------------------------------------------------------------- generator -------------

def test_generator_moments():
    u, nrm = [], []
    for i in range(20000):
        a, b = O.draw(0, 0, i % 97, i)
        u.append(a)
        nrm.append(b)
    u, nrm = np.array(u), np.array(nrm)
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(nrm.mean()) < 0.02
    assert abs(nrm.std() - 1.0) < 0.02


def test_full_step_conserves_count_and_charge():
    u = O.Universe(4, dt=0.1)
    u.init_species(0, 1, 8, -0.01, -1.0, (0.05, 0.05, 0.05))
    u.init_species(1, 1, 8, 0.01, 1.0 / 64, (0.006, 0.006, 0.006))
    q0 = u.total_charge()
    u.step(5)
    assert u.num_particles() == 2 * 8 * 64
    assert u.total_charge() == pytest.approx(q0, abs=1e-15)


# ------------------------------------------------------- mover as shipped (dipole) -------

def test_dipole_mover_field_is_the_initial_dipole_field():
    """The shipped mover's B (cell.h:919-945) is the Bext of getDipoleFieldAt (field.h:43-78,
    pinned by the field_test KATs above) for a planet at the origin, outside its radius; with
    E = 0 one oracle move equals sub-cycled Boris steps (the Boris KAT pins those) in it."""
    R, B1z, dt = 0.5, 20.0, 0.15
    pos = np.array([1.1, -0.7, 0.9])
    vel = np.array([0.3, 0.2, -0.4])
    E, Bext, B = np.zeros(3), np.zeros(3), np.zeros(3)
    O.lib().or_dipole_field(O._p(O.f64(pos)), O._p(O.f64([0, 0, 0])), O._p(O.f64([0, 0, 0])),
                            O._p(O.f64([0, 0, 0])), R, B1z, O._p(E), O._p(Bext), O._p(B))
    qom = -25.0
    n = (4, 4, 4)
    u = O.Universe(n, w=(1.0, 1.0, 1.0), origin=(-2.0, -2.0, -2.0), dt=dt, mover="dipole",
                   dipole_moment=-B1z * R ** 3)
    u.add_particles(O.make_particles(pos, vel, -1.0, qom))
    u.move()
    got, _ = u.particles()
    u.close()
    b2 = float(np.dot(Bext, Bext))
    dt_sub = np.pi * 1.0 / (4.0 * abs(qom) * b2)
    sub = min(int(dt / dt_sub) + 1, 100)
    assert sub > 1  # the sub-cycling path is exercised
    p = O.make_particles(pos, vel, -1.0, qom)
    for _ in range(sub):
        O.lib().or_update_velocity(O._p(p), O._p(O.f64([0, 0, 0])), O._p(O.f64(Bext)), dt / sub)
        O.lib().or_update_position(O._p(p), dt / sub)
    np.testing.assert_allclose(got["pos"][0], p["pos"][0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(got["vel"][0], p["vel"][0], rtol=0, atol=1e-14)


def test_static_field_policy_leaves_fields():
    """field_substeps = 0 is solveFieldStatically (field.h:272-284): E and Bc stay as set."""
    n = (4, 4, 4)
    rng = np.random.default_rng(5)
    E = rng.normal(size=(64, 3))
    Bc = rng.normal(size=(64, 3))
    u = O.Universe(n, dt=0.1, field_substeps=0)
    u.set_fields(E, Bc)
    u.set_J(rng.normal(size=(64, 3)))
    u.field_solve()
    f = u.fields()
    u.close()
    np.testing.assert_array_equal(f["E"], E)
    np.testing.assert_array_equal(f["Bc"], Bc)


# -------------------------------------------------------------- forward solver -----------

def test_forward_solver_kat():
    """field_test.cc:418-462 (solveFieldForward): uniform E 0.2, Bc 0.8, J 0.6, dt 1 ->
    E = -0.4, Bc unchanged."""
    n = (5, 5, 5)
    u = O.Universe(n, dt=1.0, solver="forward")
    nc = 125
    u.set_fields(np.full((nc, 3), 0.2), np.full((nc, 3), 0.8))
    u.set_J(np.full((nc, 3), 0.6))
    u.field_solve()
    f = u.fields()
    u.close()
    np.testing.assert_allclose(f["E"], -0.4, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f["Bc"], 0.8, rtol=0, atol=1e-15)


def test_forward_solver_speed_of_light_factors():
    """The forward solver's curls carry one factor c each (field.h:308, 322): a periodic linear
    ramp Bc_z = a*y (a jump at the wrap) gives dE_x = c dt a / w_y at interior nodes, and
    E_y = b*z gives dBc_x = +c dt b / w_z; leapfrog gives c^2 and 1 instead."""
    n, w, dt, c, a, b = (4, 6, 6), (1.0, 0.5, 2.0), 0.01, 3.0, 0.7, 0.3
    x, y, z = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    nc = np.prod(n)
    for solver, eF, bF in (("forward", c, c), ("leapfrog", c * c, 1.0)):
        E = np.zeros((nc, 3))
        Bc = np.zeros((nc, 3))
        Bc[:, 2] = (a * y).reshape(-1)
        E[:, 1] = (b * z).reshape(-1)
        u = O.Universe(n, w=w, dt=dt, c=c, solver=solver)
        u.set_fields(E, Bc)
        u.set_J(np.zeros((nc, 3)))
        u.field_solve()
        f = u.fields()
        u.close()
        inner_y = (y >= 1).reshape(-1)
        dEx = f["E"][:, 0] - E[:, 0]
        np.testing.assert_allclose(dEx[inner_y], eF * dt * a / w[1], rtol=1e-12)
        # curl+ E _x = -(dEy/dz) with E updated first (E_y unchanged by the ramp in Bc_z)
        inner_z = (z <= n[2] - 2).reshape(-1)
        dBx = f["Bc"][:, 0] - Bc[:, 0]
        np.testing.assert_allclose(dBx[inner_z], bF * dt * b / w[2], rtol=1e-12)
