"""The reference's particle initialisation with its own random streams (host code of the
library, include/ipic_gpu.h ipic_ref_*), checked on the CPU against the oracle, whose streams
come from libstdc++'s own std::minstd_rand and std::uniform_real_distribution<double>
(oracle/ref_init.cpp), and against what the mathematics fixes for the ziggurat.

Reference: initCells (cell.h:512-654), uniform_pos_normal_speed (cell.h:405-440),
ziggurat_normal_distribution.h:8-127, driver set-up src/ipic3d.cxx:41-85."""
import math
import os

import numpy as np
import pytest

import allscale_ipic3d_amd as A
from allscale_ipic3d_amd import driver as D
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _desc(n, w, origin):
    d = A.Desc()
    d.n[:] = list(n)
    d.cell_width[:] = list(w)
    d.origin[:] = list(origin)
    d.nspecies = 1
    return d


def _oracle_universe(n, w, origin):
    return O.Universe(n, w=w, origin=origin, dt=0.1, c=1.0)


def _compare(n, w, origin, lo, hi, mean, sd, num, seed=0, nthreads=0):
    d = _desc(n, w, origin)
    dist = A.ref_distribution(lo, hi, mean, sd, num, seed=seed, nthreads=nthreads)
    counts, pos, vel = A.reference_init(d, dist)
    u = _oracle_universe(n, w, origin)
    ocounts, parts = O.ref_init(u, lo, hi, mean, sd, num, 1.0, 1.0, seed=seed)
    np.testing.assert_array_equal(counts, ocounts)
    np.testing.assert_array_equal(pos, parts["pos"])
    np.testing.assert_array_equal(vel, parts["vel"])
    return counts, pos, vel


def test_driver_case_streams_bit_exact():
    """The driver's test.inp case: 8^3 cells of 1.25 Earth radii, 262,144 particles."""
    case = D.driver_case(D.Parameters(os.path.join(GOLD, "test_case.inp")), "test")
    counts, pos, vel = _compare(case.n, case.cell_width, case.origin, case.pos_min, case.pos_max, (0, 0, 0),
                                (case.v_mod,) * 3, case.num_particles)
    assert counts.sum() == case.num_particles == 262144
    # phase 1: 5 particles per pseudo particle, the remainder 6144 spread as 12 per cell; the
    # shell of cells outside the middle 75% of the box receives no pseudo particle
    assert np.all((counts - 12) % 5 == 0)
    c3 = counts.reshape(8, 8, 8)
    assert np.all(c3[0] == 12) and np.all(c3[:, :, 7] == 12) and c3[1:7, 1:7, 1:7].min() > 12
    # every particle lies in its cell, [low, low + width)
    cell = np.repeat(np.arange(512), counts)
    idx = np.stack([cell // 64, (cell // 8) % 8, cell % 8], axis=1)
    low = np.asarray(case.origin) + idx * np.asarray(case.cell_width)
    assert np.all(pos >= low) and np.all(pos < low + np.asarray(case.cell_width))


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_thread_count_does_not_change_streams(nthreads):
    n, w, o = (5, 4, 3), (1.0, 2.0, 0.5), (-1.0, 0.5, 2.0)
    lo = (o[0] + 0.5, o[1] + 1.0, o[2] + 0.2)
    hi = (o[0] + 4.0, o[1] + 7.0, o[2] + 1.4)
    _compare(n, w, o, lo, hi, (0.1, -0.2, 0.3), (1.0, 0.5, 2.0), 60 * 100 * 3 + 17, seed=12345, nthreads=nthreads)


def test_redraws_outside_the_universe():
    """Pseudo particles drawn outside the universe are redrawn (cell.h:532): a position box that
    overhangs the box on every side forces the sequential continuation from the first redraw."""
    n, w, o = (4, 4, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    counts, _, _ = _compare(n, w, o, (-1.0, -0.5, -2.0), (5.0, 4.5, 6.0), (0, 0, 0), (1, 1, 1), 64 * 100 * 2 + 5,
                            seed=7, nthreads=4)
    assert counts.sum() == 64 * 100 * 2 + 5


def test_reference_assertion_is_an_error():
    """initCells asserts that some particles remain to be spread (cell.h:557)."""
    d = _desc((2, 2, 2), (1, 1, 1), (0, 0, 0))
    dist = A.ref_distribution((0, 0, 0), (2, 2, 2), (0, 0, 0), (1, 1, 1), 8 * 100 * 3)
    with pytest.raises(A.IpicError, match="initCells"):
        A.reference_init(d, dist)


def test_uniform_stream_is_libstdcxx():
    """vector::uniform over one cell: the library's positions equal libstdc++'s
    uniform_real_distribution<double> on minstd_rand draw for draw (x, y, z interleaved)."""
    n, w, o = (1, 1, 1), (3.0, 5.0, 7.0), (1.0, -2.0, 0.5)
    d = _desc(n, w, o)
    num = 100 + 37
    dist = A.ref_distribution(o, (o[0] + 3, o[1] + 5, o[2] + 7), (0, 0, 0), (1, 1, 1), num)
    counts, pos, _ = A.reference_init(d, dist)
    u = O.ref_uniform_draws(0, 0.0, 1.0, 3 * num).reshape(num, 3)  # cell (0,0,0): seed 0
    for k in range(3):
        np.testing.assert_array_equal(pos[:, k], u[:, k] * ((o[k] + w[k]) - o[k]) + o[k])


def test_ziggurat_is_a_standard_normal():
    """The ziggurat draws (oracle restatement, which the library's velocities equal bit for bit)
    have the moments and tail mass of N(0, 1): mean 0, variance 1, fourth moment 3, and
    P(|z| > 3.44262) = erfc(3.44262 / sqrt 2) for the tail the base strip samples."""
    z = O.ref_normal_draws(2, 2_000_000).astype(np.float64)
    assert abs(z.mean()) < 4e-3
    assert abs(z.var() - 1.0) < 4e-3
    assert abs(np.mean(z ** 4) - 3.0) < 3e-2
    tail = np.mean(np.abs(z) > 3.442620)
    expect = math.erfc(3.442620 / math.sqrt(2.0))
    assert abs(tail - expect) < 5 * math.sqrt(expect / z.size)
    # a wrong level table (for example the strips' widths off by one index) shows here
    hist, edges = np.histogram(z, bins=40, range=(-4, 4))
    cdf = np.array([0.5 * math.erfc(-e / math.sqrt(2.0)) for e in edges])
    expected = np.diff(cdf) * z.size
    assert np.max(np.abs(hist - expected) / np.sqrt(expected + 1)) < 6


def test_velocities_are_mean_plus_stddev_times_the_cell_stream():
    """Cell (x,y,z) draws its velocities from a ziggurat seeded s2 = (1023 s1) & s1 with
    s1 = (1023 s) & s and s = ((x 1023) + y) 1023 + z (cell.h:278-284, 629-642)."""
    n, w, o = (2, 3, 2), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    d = _desc(n, w, o)
    num = 12 * 100 + 29
    mean, sd = (0.5, -1.0, 2.0), (3.0, 0.25, 1.5)
    dist = A.ref_distribution(o, (2, 3, 2), mean, sd, num)
    counts, _, vel = A.reference_init(d, dist)
    start = np.concatenate([[0], np.cumsum(counts)])
    for c in range(12):
        x, y, z = c // 6, (c // 2) % 3, c % 2
        s = ((x * 1023) + y) * 1023 + z
        s1 = ((1023 * s) & 0xFFFFFFFF) & s
        s2 = ((1023 * s1) & 0xFFFFFFFF) & s1
        k = int(counts[c])
        zz = O.ref_normal_draws(s2, 3 * k).astype(np.float64).reshape(k, 3)
        np.testing.assert_array_equal(vel[start[c]:start[c + 1]], np.asarray(mean) + np.asarray(sd) * zz)
