"""The benchmark's workloads at their full sizes against the FP64 oracle on the host cores.

Both runs start from the same particles (the GPU's generator, FP32 values widened to FP64)
and the same fields.  The particle order inside a cell differs, so particles are compared
through per-cell sums (oracle/compare.py): for a cell with the same count, the per-cell sums of
offsets and velocities differ by at most the sum of the per-particle differences, so

  * offset_max   = max over cells |sum_gpu(offset) - sum_ref(offset)| / count   <= 1e-5 cells
  * velocity_max = max over cells |sum_gpu(v) - sum_ref(v)| / sum_ref(|v|)     <= 1e-5

bound the FP32-vs-FP64 particle error by the north-star tolerance.  Fields (E, Bc, Bn, J) are
compared in relative l2 norm (<= 1e-5), the FP64 conservation sums relatively (<= 1e-5).
Cells whose counts differ (particles that the FP32 and FP64 runs place on different sides of a
face) are reported and bounded.  Every checkpoint prints its errors, so a run shows how they
grow with the step count."""
import json

import numpy as np
import pytest

import oracle.compare as OC

pytestmark = pytest.mark.gpu

TOL = 1e-5
FOUR_PI = 16.0 * np.arctan(1.0)


def _two_species_case(n, ppc, seed):
    import allscale_ipic3d_amd as G
    half = ppc // 2
    q = 1.0 / (half * FOUR_PI)
    species = ((-q, -1.0), (q, 1.0 / 64))
    sim = G.Simulation(n, dt=0.05, c=1.0, species=species)
    sim.init_particles(0, half, (0.05,) * 3, seed=seed)
    sim.init_particles(1, ppc - half, (0.006,) * 3, seed=seed)
    rng = np.random.default_rng(seed)
    E = rng.normal(0.0, 1e-3, size=(n ** 3, 3))
    Bc = np.tile([0.0, 0.2, 1.0], (n ** 3, 1))
    sim.upload_fields(E, Bc)
    u = OC.oracle_from_gpu(sim, species, E, Bc, dt=0.05, c=1.0, field_substeps=1)
    return sim, u, species


def _check(res, step, npart, max_flip_frac):
    print(f"step {step}:", json.dumps(res), flush=True)
    for k, e in res["fields"].items():
        assert e <= TOL, f"step {step}: {k} rel error {e:.3e}"
    for s, p in enumerate(res["particles"]):
        assert p["particles"] == p["particles_ref"]
        assert p["offset_max"] <= TOL, f"step {step} species {s}: offset {p['offset_max']:.3e}"
        assert p["velocity_max"] <= TOL, f"step {step} species {s}: velocity {p['velocity_max']:.3e}"
        assert p["flip_particles"] <= max_flip_frac * npart, f"step {step} species {s}: {p['flip_particles']} flips"
    c = res["conservation"]
    assert c["particles"] == 0
    for k in ("kinetic_energy", "momentum", "field_energy_E", "field_energy_B"):
        assert c[k] <= TOL, f"step {step}: {k} rel error {c[k]:.3e}"
    assert c["charge"] <= 1e-12


@pytest.mark.timeout(1200)
def test_config5_128c_32ppc_against_oracle():
    """128^3 cells x 32 particles per cell (67,108,864), two species: the benchmark's workload,
    the timed kernel instance, 5 steps."""
    n, ppc = 128, 32
    sim, u, species = _two_species_case(n, ppc, seed=1)
    npart = n ** 3 * ppc
    done = 0
    for step in (1, 5):
        u.step(step - done)
        sim.step(step - done)
        done = step
        _check(OC.compare(u, sim, species), step, npart, 1e-5)
    sim.close()
    u.close()


@pytest.mark.timeout(1200)
def test_config2_64c_64ppc_100_steps_against_oracle():
    """64^3 cells x 64 particles per cell (16,777,216), two species, the configuration's 100
    steps; errors at steps 1, 10, 25, 50 and 100."""
    n, ppc = 64, 64
    sim, u, species = _two_species_case(n, ppc, seed=2)
    npart = n ** 3 * ppc
    done = 0
    for step in (1, 10, 25, 50, 100):
        u.step(step - done)
        sim.step(step - done)
        done = step
        _check(OC.compare(u, sim, species), step, npart, 1e-4)
    sim.close()
    u.close()
