"""Edge cases of the GPU step through the C ABI: empty universes and empty species, a species
that empties during the run, refused calls that leave the context usable, and the per-species
size limit."""
import numpy as np
import pytest

import allscale_ipic3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def test_empty_universe_steps():
    """No particles at all: the step runs the field phase alone; counts and sums stay zero and
    the fields evolve as the oracle's leapfrog does with J = 0."""
    n = (8, 8, 8)
    rng = np.random.default_rng(1)
    E = rng.normal(0.0, 1e-2, size=(512, 3))
    Bc = rng.normal(0.0, 1e-2, size=(512, 3)) + np.array([0.0, 0.0, 1.0])
    sim = G.Simulation(n, dt=0.1, species=((-1.0, -1.0), (1.0, 1.0)))
    sim.upload_particles(0, np.zeros((0, 3)), np.zeros((0, 3)))
    sim.upload_fields(E, Bc)
    st = sim.step(3)
    assert st["particles"] == 0 and st["removed"] == 0
    assert sim.num_particles() == 0
    assert not sim.cell_counts().any()
    c = sim.conservation()
    assert c["particles"] == 0 and c["kinetic_energy"] == 0.0
    u = O.Universe(n, dt=0.1)
    u.set_fields(E, Bc)
    u.step(3)
    ref, got = u.fields(), sim.download_fields()
    for k in ("E", "Bc"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-5, atol=1e-7)
    sim.close()
    u.close()


def test_one_species_empty():
    """Species 1 has no particles; species 0 moves as in the oracle."""
    n = (6, 4, 5)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0.0, 1.0, size=(400, 3)) * np.asarray(n)
    vel = rng.normal(0.0, 0.3, size=(400, 3))
    sim = G.Simulation(n, dt=0.1, species=((-0.01, -1.0), (0.01, 1.0)))
    sim.upload_particles(0, pos, vel)
    sim.upload_fields(np.zeros((120, 3)), np.tile([0.0, 0.0, 0.5], (120, 1)))
    sim.step(5)
    u = O.Universe(n, dt=0.1)
    u.add_particles(O.make_particles(pos, vel, -0.01, -1.0))
    u.set_fields(np.zeros((120, 3)), np.tile([0.0, 0.0, 0.5], (120, 1)))
    u.step(5)
    assert sim.num_particles(1) == 0 and sim.num_particles(0) == 400
    np.testing.assert_array_equal(sim.cell_counts(0), u.cell_counts())
    sim.close()
    u.close()


def test_species_empties_through_the_sink():
    """Every particle starts inside the sink sphere: all are removed in the first step and the
    following steps run on an empty species."""
    n = (8, 8, 8)
    rng = np.random.default_rng(3)
    pos = np.array([4.0, 4.0, 4.0]) + rng.uniform(-0.5, 0.5, size=(300, 3))
    vel = rng.normal(0.0, 0.01, size=(300, 3))
    sim = G.Simulation(n, dt=0.1, species=((-0.01, -1.0),), sink_center=(4.0, 4.0, 4.0), sink_radius=2.0)
    sim.upload_particles(0, pos, vel)
    st = sim.step(1)
    assert st["removed"] == 300 and sim.num_particles() == 0
    st = sim.step(3)
    assert st["removed"] == 0 and sim.num_particles() == 0
    sim.close()


def test_refused_calls_leave_the_context_usable():
    n = (4, 4, 4)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0.0, 4.0, size=(100, 3))
    vel = rng.normal(0.0, 0.1, size=(100, 3))
    sim = G.Simulation(n, dt=0.1, species=((-0.01, -1.0),))
    sim.upload_particles(0, pos, vel)
    with pytest.raises(G.IpicError, match="species out of range"):
        sim.upload_particles(3, pos, vel)
    with pytest.raises(G.IpicError, match="non-finite"):
        bad = pos.copy()
        bad[7, 1] = np.nan
        sim.upload_particles(0, bad, vel)
    # the per-species limit (one buffer resource spans the six arrays, < 4 GiB), refused before
    # any allocation
    big = G.Simulation((128, 128, 128), dt=0.1, species=((-0.01, -1.0),))
    with pytest.raises(G.IpicError, match="178956970"):
        big.init_particles(0, 86, (0.1, 0.1, 0.1))
    # the same limit on upload, checked before the host builds any buffer (the arrays passed
    # are never read: the call is refused first)
    tiny = np.zeros((1, 3))
    rc = big._L.ipic_gpu_upload_particles(big._h, 0, 178956971, G._ptr(tiny), G._ptr(tiny))
    assert rc == -4 and "178956970" in big._L.ipic_gpu_last_error(big._h).decode()
    big.init_particles(0, 2, (0.1, 0.1, 0.1))
    assert big.step(1)["particles"] == 2 * 128 ** 3
    big.close()
    # the first context still holds its particles and steps
    st = sim.step(2)
    assert st["particles"] == 100 and sim.num_particles() == 100
    sim.close()
