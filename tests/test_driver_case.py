"""The reference's main driver case (src/ipic3d.cxx:41-99 on inputs/test.inp and tiny.inp) in the
FP64 oracle: set-up, the reference's initial streams, 30 (100) steps of the loop as shipped
(simulator.h:181-231: dipole mover, reflection, planet sink), against the reference's own
outputs (outputs/test.out, tiny.out and the ConservedQuantities logs, kept under tests/golden/).

What is pinned exactly: the derived set-up, the initial field energies of the logs (every
printed digit, both cases), the particle count.  The initial particles are the ones the
reference's code draws (bit-exact against libstdc++'s generators, test_reference_init.py), yet
the reference's recorded run does not start from exactly this state: its logged initial kinetic
energy and momentum are 0.26% and 0.14% below these particles' (the tiny log even repeats the
test log's particle sums for a run with eight times the particles), and its final counts differ
from the oracle's: 0.7% (test) and 0.65% (tiny) of the particles sit in another cell, against
several times that for an independent draw of the same distribution.  The counts agree best at exactly
the recorded number of steps (the distance grows by 20-90% one step either side), so the
statistics of the recorded run and its dynamics are reproduced; bit-exact counts are not
("parity partial" for the end-to-end outputs)."""
import math
import os

import numpy as np
import pytest

from allscale_ipic3d_amd import driver as D, io
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case(name):
    return D.driver_case(D.Parameters(os.path.join(GOLD, f"{name}_case.inp")), name)


def _gold_rows(name):
    with open(os.path.join(GOLD, f"{name}_conserved.txt")) as f:
        rows = [r.split("\t") for r in f.read().splitlines()[1:] if r.strip()]
    return [(int(r[0]), float(r[1]), float(r[2]), float(r[3]), float(r[4])) for r in rows]


def _oracle(case):
    u = O.Universe(case.n, w=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                   sink_center=case.object_center, sink_radius=case.planet_radius, bc="reflect", mover="dipole",
                   dipole_moment=case.dipole_moment, field_substeps=0)
    counts, parts = O.ref_init(u, case.pos_min, case.pos_max, (0, 0, 0), (case.v_mod,) * 3, case.num_particles,
                               *case.species)
    u.add_particles(parts)
    return u, counts


def _sums(u, mass):
    p, _ = u.particles()
    v2 = (p["vel"] ** 2).sum(1)
    return float((mass * np.sqrt(v2)).sum()), float(0.5 * mass * v2.sum())


def test_driver_setup():
    """ipic3d.cxx:41-85 on test.inp: 8^3 cells of 10 / 8 Earth radii, the box centred on the
    planet, 8 x 8 x 8 particles per cell, the speed of a 10 MeV proton."""
    c = _case("test")
    R = 6378137.0
    assert c.n == (8, 8, 8) and c.ncycles == 30 and c.field_output_cycle == 1
    assert c.cell_width == (1.25 * R,) * 3
    assert c.origin == (-5.0 * R,) * 3
    assert c.num_particles == 262144
    assert c.pos_min == (-3.75 * R,) * 3 and c.pos_max == (3.75 * R,) * 3
    # the speed of ipic3d.cxx:74 for K = 10 MeV: close to sqrt(K / m), 3.08e7 m/s
    assert abs(c.v_mod - 30786068.314846) < 1e-5
    assert c.dipole_moment == -3.07e-5 * R ** 3
    t = _case("tiny")
    assert t.n == (16, 16, 16) and t.ncycles == 100 and t.field_output_cycle == 4
    assert t.num_particles == 16 ** 3 * 512


@pytest.mark.parametrize("name", ["test", "tiny"])
def test_initial_field_energies_match_the_log(name):
    """The E and B (+ Bext) energies of the reference's logs, every printed digit."""
    e, b = D.initial_field_energies(_case(name))
    row = _gold_rows(name)[0]
    assert f"{e:g}" == f"{row[2]:g}"
    assert f"{b:g}" == f"{row[3]:g}"


def _distance(a, b):
    return int(np.abs(np.asarray(a) - np.asarray(b)).sum())


@pytest.mark.parametrize("name", ["test", "tiny"])
def test_oracle_reproduces_reference_outputs_statistically(name):
    case = _case(name)
    _, gold = io.read_cell_counts(os.path.join(GOLD, f"{name}_out_counts.txt"))
    rows = _gold_rows(name)
    u, counts = _oracle(case)
    assert counts.sum() == case.num_particles
    mass = case.species[0] / case.species[1]
    every = case.field_output_cycle
    dist = {}
    k = 0
    for i in range(case.ncycles + 2):
        if i < case.ncycles and (i % every == 0 or i + 1 == case.ncycles):
            mom, ke = _sums(u, mass)
            g = rows[k]
            k += 1
            assert g[0] == i
            # the logged particle sums track these within the offset of the initial state (the
            # tiny log's particle columns are the test run's, test_tiny_log_repeats_test_sums)
            if name == "test":
                assert abs(mom / g[1] - 1) < 2.5e-3 and abs(ke / g[4] - 1) < 4e-3
        if i >= case.ncycles - 1:
            dist[i] = _distance(u.cell_counts(), gold)
        if i == case.ncycles:
            n = u.cell_counts().sum()
        if i <= case.ncycles:
            u.move(8)
            u.rebin(8)
    assert k == len(rows)
    print(name, "total", n, "reference", gold.sum(), "L1", dist)
    assert abs(n / gold.sum() - 1) < 5e-4
    # particles in another cell than the reference's: half the L1 distance of the counts
    assert dist[case.ncycles] / 2 <= 0.01 * gold.sum()
    # closest at exactly the recorded number of steps
    assert dist[case.ncycles] < 0.9 * min(dist[case.ncycles - 1], dist[case.ncycles + 1])
    u.close()


def test_tiny_log_repeats_test_sums():
    """The reference's tiny log holds, at every cycle both logs record, the particle momentum and
    kinetic energy of the test log, though its run has eight times the particles."""
    t = {r[0]: r for r in _gold_rows("test")}
    common = [r for r in _gold_rows("tiny") if r[0] in t]
    assert len(common) >= 8
    for r in common:
        assert r[1] == t[r[0]][1] and r[4] == t[r[0]][4]


def test_independent_draw_is_farther_from_the_reference():
    """Calibration of the statistical bar: the same case drawn with another distribution seed
    ends 2.5x or more farther from outputs/test.out than the reference's own streams."""
    case = _case("test")
    _, gold = io.read_cell_counts(os.path.join(GOLD, "test_out_counts.txt"))
    u = O.Universe(case.n, w=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                   sink_center=case.object_center, sink_radius=case.planet_radius, bc="reflect", mover="dipole",
                   dipole_moment=case.dipole_moment, field_substeps=0)
    _, parts = O.ref_init(u, case.pos_min, case.pos_max, (0, 0, 0), (case.v_mod,) * 3, case.num_particles,
                          *case.species, seed=1000)
    u.add_particles(parts)
    for _ in range(case.ncycles):
        u.move(8)
        u.rebin(8)
    far = _distance(u.cell_counts(), gold)
    u2, _ = _oracle(case)
    for _ in range(case.ncycles):
        u2.move(8)
        u2.rebin(8)
    near = _distance(u2.cell_counts(), gold)
    print("own streams", near, "other seed", far)
    assert far > 2.5 * near
