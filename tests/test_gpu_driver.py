"""The reference's main driver case (src/ipic3d.cxx on inputs/test.inp) on the GPU: the
library's reference-stream initialisation (ipic_gpu_init_reference) and 30 steps of the loop as
shipped (analytic dipole mover, reflection, planet sink; FP32 on the device) against the FP64
oracle run of the same case, step by step, and against the reference's outputs/test.out."""
import os

import numpy as np
import pytest

import allscale_ipic3d_amd as A
from allscale_ipic3d_amd import driver as D, io
import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case():
    return D.driver_case(D.Parameters(os.path.join(GOLD, "test_case.inp")), "test")


def test_driver_case_matches_oracle_and_reference(tmp_path):
    case = _case()
    sim = D.make_simulation(case)
    counts = sim.init_reference(0, D.case_distribution(case))
    u = O.Universe(case.n, w=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                   sink_center=case.object_center, sink_radius=case.planet_radius, bc="reflect", mover="dipole",
                   dipole_moment=case.dipole_moment, field_substeps=0)
    ocounts, parts = O.ref_init(u, case.pos_min, case.pos_max, (0, 0, 0), (case.v_mod,) * 3, case.num_particles,
                                *case.species)
    u.add_particles(parts)
    np.testing.assert_array_equal(counts, ocounts)
    np.testing.assert_array_equal(sim.cell_counts(), u.cell_counts())
    p = sim.download_particles(0)
    # the device holds FP32 offsets in the cell and FP32 velocities of the same particles, in
    # the same order (row-major cells, draw order within a cell)
    np.testing.assert_allclose(p["pos"], parts["pos"], rtol=0, atol=2.0)  # metres, cells of 8e6 m
    np.testing.assert_allclose(p["vel"], parts["vel"], rtol=1e-7, atol=0)
    mass = case.species[0] / case.species[1]
    worst = 0
    for i in range(case.ncycles):
        c = sim.conservation()
        op, _ = u.particles()
        v2 = (op["vel"] ** 2).sum(1)
        assert abs(c["kinetic_energy"] / (0.5 * mass * v2.sum()) - 1) < 1e-5
        assert abs(c["momentum"] / (mass * np.sqrt(v2)).sum() - 1) < 1e-5
        sim.step(1)
        u.move(8)
        u.rebin(8)
        gc, oc = sim.cell_counts(), u.cell_counts()
        worst = max(worst, int(np.abs(gc - oc).sum()))
        print(f"step {i + 1}: GPU {gc.sum()} oracle {oc.sum()} L1 {int(np.abs(gc - oc).sum())}", flush=True)
    # FP32 against FP64 from the same state: a particle's cell may differ where it ends within
    # the FP32 rounding of a face; at most 1e-4 of the particles
    assert worst / 2 <= 1e-4 * case.num_particles
    _, gold = io.read_cell_counts(os.path.join(GOLD, "test_out_counts.txt"))
    final = sim.cell_counts()
    assert abs(final.sum() / gold.sum() - 1) < 5e-4
    assert np.abs(final - gold).sum() / 2 <= 0.01 * gold.sum()
    sim.close()
    u.close()


def test_driver_run_writes_reference_formats(tmp_path):
    res = D.run(os.path.join(GOLD, "test_case.inp"), out_dir=str(tmp_path))
    n, counts = io.read_cell_counts(str(tmp_path / "test_case.out"))
    assert n == (8, 8, 8)
    np.testing.assert_array_equal(counts, res["counts"])
    with open(tmp_path / "test_caseConservedQuantities.out") as f:
        rows = f.read().splitlines()
    with open(os.path.join(GOLD, "test_conserved.txt")) as f:
        gold = f.read().splitlines()
    assert rows[0] == gold[0] and len(rows) == len(gold)
    for r, g in zip(rows[1:], gold[1:]):
        r, g = r.split("\t"), g.split("\t")
        assert r[0] == g[0] and r[2] == g[2] and r[3] == g[3]   # cycle, E and B energies


def test_tiny_case_matches_oracle_and_reference():
    """inputs/tiny.inp: 16^3 cells x 512 particles (2,097,152), 100 steps, against the oracle
    every 25 steps and against outputs/tiny.out at the end."""
    case = D.driver_case(D.Parameters(os.path.join(GOLD, "tiny_case.inp")), "tiny")
    sim = D.make_simulation(case)
    counts = sim.init_reference(0, D.case_distribution(case))
    u = O.Universe(case.n, w=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                   sink_center=case.object_center, sink_radius=case.planet_radius, bc="reflect", mover="dipole",
                   dipole_moment=case.dipole_moment, field_substeps=0)
    ocounts, parts = O.ref_init(u, case.pos_min, case.pos_max, (0, 0, 0), (case.v_mod,) * 3, case.num_particles,
                                *case.species)
    u.add_particles(parts)
    np.testing.assert_array_equal(counts, ocounts)
    for block in range(4):
        sim.step(25)
        for _ in range(25):
            u.move(0)
            u.rebin(0)
        gc, oc = sim.cell_counts(), u.cell_counts()
        moved = int(np.abs(gc - oc).sum()) // 2
        print(f"step {25 * (block + 1)}: GPU {gc.sum()} oracle {oc.sum()} in another cell {moved}", flush=True)
        assert moved <= 1e-4 * case.num_particles
    _, gold = io.read_cell_counts(os.path.join(GOLD, "tiny_out_counts.txt"))
    final = sim.cell_counts()
    assert abs(final.sum() / gold.sum() - 1) < 5e-4
    assert np.abs(final - gold).sum() / 2 <= 0.01 * gold.sum()
    sim.close()
    u.close()


@pytest.mark.parametrize("case", ["test", "tiny"])
def test_lattice_init_on_device_matches_the_logged_initial_state(case):
    """The parameter-file initCells (cell.h:699-776) uploaded by ipic_gpu_init_lattice: the
    device holds the host lattice particle for particle, and its conservation sums and the field
    energies of ipic_gpu_init_fields give row 0 of outputs/{test,tiny}ConservedQuantities.static.out
    (8^3 x 5^3 and 16^3 x 8^3 particles, qom -25, uth 0.045, u0 0.02, B0 = (0, 0, 1e-4))."""
    n, w, ppc, gold = {"test": (8, 1.25, 5, "test_conserved_static.txt"),
                       "tiny": (16, 0.625, 8, "tiny_conserved_static.txt")}[case]
    with open(os.path.join(GOLD, gold)) as f:
        mom_log, e_log, b_log, ke_log = (float(v) for v in f.read().splitlines()[1].split()[1:])
    q = A.lattice_charge((w,) * 3, (ppc,) * 3, -25.0)
    sim = A.Simulation((n,) * 3, cell_width=(w,) * 3, origin=(0.0, 0.0, 0.0), dt=0.15, c=1.0,
                       species=((q, -25.0),), field_substeps=0, flags=A.FLAG_REFLECT)
    lat = A.ref_lattice((ppc,) * 3, (0.045,) * 3, (0.02, 0.0, 0.0))
    sim.init_lattice(0, lat)
    e_en, b_en = sim.init_fields(A.field_init(drift=(0.02, 0.0, 0.0), B0=(0.0, 0.0, 1e-4)))
    d = A.Desc()
    d.n[:] = [n] * 3
    d.cell_width[:] = [w] * 3
    d.nspecies = 1
    pos, vel = A.lattice_particles(d, lat)
    p = sim.download_particles(0)
    np.testing.assert_array_equal(sim.cell_counts(), np.full(n ** 3, ppc ** 3))
    np.testing.assert_allclose(p["pos"], pos, rtol=0, atol=1e-6 * w)    # FP32 offsets in the cell
    np.testing.assert_allclose(p["vel"], vel, rtol=1e-7, atol=1e-12)
    c = sim.conservation()
    assert abs(c["momentum"] / mom_log - 1) < 1e-5
    assert abs(c["kinetic_energy"] / ke_log - 1) < 1e-5
    assert abs(e_en / e_log - 1) < 5e-6
    assert b_en <= b_log and abs(b_en / b_log - 1) < 1e-3
    sim.close()
