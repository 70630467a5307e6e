"""Reproducible steps (IPIC_FLAG_REPRODUCIBLE).  After every pass that moves particles the
library puts each cell's arrivals (the particles that changed cell, which take their slots by
atomic counters) in the order of their own contents, so the state after a step is a function of
the state before it: two runs from one seed agree bit for bit, fields, conservation sums,
counts, particles and ids.  Without the flag they agree to FP32 rounding only
(test_gpu_determinism.py).  The reference imports arrivals sequentially (cell.h:1272-1306), a
fixed order as well, though not this one; the ordered path is compared with the FP64 oracle at
the same bars as the default path."""
import numpy as np
import pytest

import allscale_ipic3d_amd as G
import oracle as O
from tests.test_gpu_parity import _compare_fields, _compare_particles

pytestmark = pytest.mark.gpu


def _snapshot(sim, nsp):
    return {"fields": sim.download_fields(), "cons": sim.conservation(),
            "counts": [np.asarray(sim.cell_counts(s)) for s in range(nsp)],
            "parts": [sim.download_particles(s) for s in range(nsp)]}


def _assert_bitwise_equal(a, b):
    for key, v in a["fields"].items():
        assert np.array_equal(v, b["fields"][key]), key
    for key, v in a["cons"].items():
        w = b["cons"][key]
        assert v == w or (v != v and w != w), key
    for ca, cb in zip(a["counts"], b["counts"]):
        assert np.array_equal(ca, cb)
    for pa, pb in zip(a["parts"], b["parts"]):
        assert pa.keys() == pb.keys()
        for key in pa:
            assert np.array_equal(pa[key], pb[key]), key


def _hot_run(track_ids):
    # the configuration of test_gpu_determinism.py: 32^3 cells x 32 ppc, two species, hot
    # electrons (about one arrival per cell and step), 12 steps
    n, ppc, seed = 32, 32, 5
    half = ppc // 2
    q = 1.0 / (half * 16.0 * np.arctan(1.0))
    sim = G.Simulation(n, dt=0.05, c=1.0, species=((-q, -1.0), (q, 1.0 / 64)), track_ids=track_ids,
                       flags=G.FLAG_REPRODUCIBLE)
    sim.init_particles(0, half, (0.4,) * 3, seed=seed)
    sim.init_particles(1, ppc - half, (0.02,) * 3, seed=seed)
    rng = np.random.default_rng(seed)
    sim.upload_fields(rng.normal(0.0, 1e-3, size=(n ** 3, 3)), np.tile([0.0, 0.2, 1.0], (n ** 3, 1)))
    sim.step(12)
    out = _snapshot(sim, 2)
    sim.close()
    return out


@pytest.mark.parametrize("track_ids", [False, True])
def test_two_runs_from_one_seed_are_bitwise_equal(track_ids):
    a, b = _hot_run(track_ids), _hot_run(track_ids)
    _assert_bitwise_equal(a, b)
    assert np.any(a["fields"]["J"] != 0.0)
    assert sum(np.count_nonzero(c != 16) for c in a["counts"]) > 0  # particles did change cell


def _crowded_case():
    # 4^3 cells x 400 particles moving up to 0.6 cells per step: every cell receives about a
    # hundred arrivals per step, so the ranges go to the one-block-per-cell kernel
    rng = np.random.default_rng(17)
    n = (4, 4, 4)
    N = 400 * 64
    pos = rng.uniform(size=(N, 3)) * np.array(n)
    vel = rng.uniform(-0.6, 0.6, size=(N, 3))
    vel[::3] *= 0.01  # and some slow ones that stay
    return n, pos, vel


def _crowded_gpu(n, pos, vel, steps):
    sim = G.Simulation(n, dt=1.0, species=((0.01, -1.0),), track_ids=True, flags=G.FLAG_REPRODUCIBLE)
    sim.upload_particles(0, pos, vel)
    sim.step(steps)
    return sim


def test_crowded_arrivals_bitwise_equal_and_match_oracle():
    n, pos, vel = _crowded_case()
    steps = 3
    a_sim = _crowded_gpu(n, pos, vel, steps)
    a = _snapshot(a_sim, 1)
    b_sim = _crowded_gpu(n, pos, vel, steps)
    b = _snapshot(b_sim, 1)
    b_sim.close()
    _assert_bitwise_equal(a, b)
    u = O.Universe(n, w=(1.0, 1.0, 1.0), dt=1.0, c=1.0)
    u.add_particles(O.make_particles(pos, vel, 0.01, -1.0))
    u.step(steps)
    _compare_fields(u, a_sim)
    _compare_particles(u, a_sim, [-1.0], n, (1, 1, 1), 0.6)
    a_sim.close()
    u.close()


def test_shipped_policies_bitwise_equal_and_match_oracle():
    """The reference's default composition (dipole mover, reflection at the global faces, planet
    sink, static field; simulator.h:244-276) with reproducible order: removed and reflected
    arrivals go through the ordering too."""
    n, origin = (8, 8, 8), (-4.0, -4.0, -4.0)
    R, B1z, dt, q, qom = 1.0, 0.5, 0.15, -0.01, -25.0
    rng = np.random.default_rng(31)
    pos = rng.uniform(-4.0, 4.0, size=(40000, 3))
    pos = pos[np.sum(pos * pos, axis=1) > (1.2 * R) ** 2]
    vel = rng.normal(0.0, 0.2, size=(len(pos), 3)) + np.array([0.05, 0.0, 0.0])

    def gpu():
        sim = G.Simulation(n, origin=origin, dt=dt, species=((q, qom),), field_substeps=0,
                           sink_center=(0.0, 0.0, 0.0), sink_radius=R, track_ids=True,
                           flags=G.FLAG_REFLECT | G.FLAG_REPRODUCIBLE, dipole_moment=-B1z * R ** 3)
        sim.upload_particles(0, pos, vel)
        st = sim.step(5)
        return sim, st

    a_sim, st = gpu()
    b_sim, _ = gpu()
    assert st["removed"] > 0
    _assert_bitwise_equal(_snapshot(a_sim, 1), _snapshot(b_sim, 1))
    b_sim.close()
    u = O.Universe(n, origin=origin, dt=dt, field_substeps=0, sink_center=(0.0, 0.0, 0.0), sink_radius=R,
                   bc="reflect", mover="dipole", dipole_moment=-B1z * R ** 3)
    u.add_particles(O.make_particles(pos, vel, q, qom))
    u.step(5)
    _compare_particles(u, a_sim, [qom], n, (1, 1, 1), 0.2)
    np.testing.assert_array_equal(a_sim.cell_counts(), u.cell_counts())
    a_sim.close()
    u.close()


def test_reference_driver_case_bitwise_equal_and_near_oracle():
    """The reference's main driver case (inputs/test.inp: 262,144 particles, 1,000+ per populated
    cell, about a hundred arrivals per cell and step) for its 30 steps, twice with reproducible
    order: the runs agree bit for bit, and the counts stay within the default path's bound of the
    FP64 oracle (a particle that FP32 and FP64 place on either side of a face)."""
    import os
    from allscale_ipic3d_amd import driver as D
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    case = D.driver_case(D.Parameters(os.path.join(gold, "test_case.inp")), "test")

    def gpu():
        sim = D.make_simulation(case, flags=G.FLAG_REPRODUCIBLE)
        sim.init_reference(0, D.case_distribution(case))
        sim.step(case.ncycles)
        return sim

    a_sim, b_sim = gpu(), gpu()
    _assert_bitwise_equal(_snapshot(a_sim, 1), _snapshot(b_sim, 1))
    b_sim.close()
    u = O.Universe(case.n, w=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                   sink_center=case.object_center, sink_radius=case.planet_radius, bc="reflect", mover="dipole",
                   dipole_moment=case.dipole_moment, field_substeps=0)
    _, parts = O.ref_init(u, case.pos_min, case.pos_max, (0, 0, 0), (case.v_mod,) * 3, case.num_particles,
                          *case.species)
    u.add_particles(parts)
    for _ in range(case.ncycles):
        u.move(8)
        u.rebin(8)
    gc, oc = a_sim.cell_counts(), u.cell_counts()
    print("L1 count difference against the oracle after", case.ncycles, "steps:", int(np.abs(gc - oc).sum()))
    assert np.abs(gc - oc).sum() / 2 <= 1e-4 * case.num_particles
    a_sim.close()
    u.close()


def test_driver_runs_write_identical_files(tmp_path):
    """The driver's --reproducible runs of the reference's test case (30 cycles) write
    byte-identical count and conserved-quantity files."""
    import os
    from allscale_ipic3d_amd import driver as D
    cfg = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_case.inp")
    for d in ("a", "b"):
        D.run(cfg, out_dir=str(tmp_path / d), reproducible=True)
    for name in ("test_case.out", "test_caseConservedQuantities.out"):
        a = (tmp_path / "a" / name).read_bytes()
        assert len(a) > 0 and a == (tmp_path / "b" / name).read_bytes(), name
