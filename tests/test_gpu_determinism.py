"""Run-to-run agreement of the GPU step.  Two simulations from the same seed and inputs (32^3
cells x 32 particles per cell, two species, hot electrons so that many particles cross cells and
tiles and deposit beyond their tile) are stepped 12 times each.

What is exact and what is not: arrivals take their output slots with atomic adds, so the order
of particles inside a cell differs from run to run.  J is summed exactly (integer fixed point in
the tile and overflow accumulators, FP64 across tiles in a fixed order), but each run of up to
five particles first adds its moments in FP32, and which particles share a run follows the
order; so J, and through it the fields and the particles, differ between runs at the level of
FP32 rounding.  The test pins that level: fields within 1e-6 of each other (relative), each
tracked particle within 1e-6 of itself, per-cell counts equal up to rare face flips, and the
conservation sums within 1e-9 (E.J 1e-8).  Measured: fields 7.5e-8 (E) and 1.4e-7 (J) apart,
particles 3e-8 (offsets) and 1.2e-7 (velocities), counts identical, kinetic energy 2e-12."""
import numpy as np
import pytest

import allscale_ipic3d_amd as G

pytestmark = pytest.mark.gpu

N, PPC, SEED, STEPS = 32, 32, 5, 12


def _run(track_ids):
    half = PPC // 2
    q = 1.0 / (half * 16.0 * np.arctan(1.0))
    sim = G.Simulation(N, dt=0.05, c=1.0, species=((-q, -1.0), (q, 1.0 / 64)), track_ids=track_ids)
    sim.init_particles(0, half, (0.4,) * 3, seed=SEED)
    sim.init_particles(1, PPC - half, (0.02,) * 3, seed=SEED)
    rng = np.random.default_rng(SEED)
    sim.upload_fields(rng.normal(0.0, 1e-3, size=(N ** 3, 3)), np.tile([0.0, 0.2, 1.0], (N ** 3, 1)))
    sim.step(STEPS)
    out = {"fields": sim.download_fields(), "cons": sim.conservation(),
           "counts": [sim.cell_counts(s) for s in range(2)],
           "parts": [sim.download_particles(s) for s in range(2)]}
    sim.close()
    return out


def _rel(a, b, ref):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.parametrize("track_ids", [False, True])
def test_two_runs_from_one_seed_agree_to_fp32_rounding(track_ids):
    a, b = _run(track_ids), _run(track_ids)
    for key in ("E", "Bc", "Bn", "J"):
        assert _rel(a["fields"][key], b["fields"][key], a["fields"][key]) <= 1e-6, key
    # the background field and B0 dominate Bc and Bn: the evolving parts agree as well
    for key in ("E", "J"):
        assert np.any(a["fields"][key] != 0.0)
    npart = N ** 3 * PPC // 2
    for s in range(2):
        ca, cb = np.asarray(a["counts"][s]), np.asarray(b["counts"][s])
        assert ca.sum() == cb.sum() == npart
        assert np.abs(ca - cb).sum() <= 1e-4 * npart
        pa, pb = a["parts"][s], b["parts"][s]
        if track_ids:
            ia, ib = np.argsort(pa["ids"]), np.argsort(pb["ids"])
            assert np.array_equal(pa["ids"][ia], pb["ids"][ib])
            assert np.max(np.abs(pa["pos"][ia] - pb["pos"][ib])) <= 1e-6 * N
            assert np.max(np.abs(pa["vel"][ia] - pb["vel"][ib])) <= 1e-6 * max(np.max(np.abs(pa["vel"])), 1e-300)
            assert np.count_nonzero(pa["cell"][ia] != pb["cell"][ib]) <= 1e-4 * npart
    # (div B is a roundoff-level residual, 1e-9 here, whose own digits are noise: not compared)
    for key, tol in (("kinetic_energy", 1e-9), ("momentum", 1e-9), ("charge", 1e-9), ("particles", 0.0),
                     ("field_energy_E", 1e-9), ("field_energy_B", 1e-9), ("EdotJ", 1e-8)):
        v, w = a["cons"][key], b["cons"][key]
        assert abs(v - w) <= tol * max(abs(v), abs(w), 1e-300), key
    moved = sum(np.count_nonzero(np.asarray(a["counts"][s]) != PPC // 2) for s in range(2))
    assert moved > 0
