"""The reference's output formats, written from a GPU-resident state.

    outputNumberOfParticlesPerCell  cell.h:1360-1395   write_cell_counts / read_cell_counts
    outputParticlePositions         cell.h:1397-1429   write_particle_positions
    writeOutputHeader / writeOutputData (conserved quantities and density frames)
                                    simulator.h:67-115 ConservedLog, write_density_frame
    getFieldEnergy                  field.h:443-459    reference_field_energy
    outputFieldGrids                field.h:464-530    write_field_grids

Numbers are printed as a default C++ ostream prints a double (6 significant digits, %g);
coordinates and vectors as "[x,y,z]" (the grid size line of outputs/test.out).  The grids of
this build are N-periodic (N^3 nodes), so the field dump has N^3 entries where the reference's
ghost-padded grids have (N+3)^3 and (N+2)^3.
"""
from __future__ import annotations

import numpy as np

FOUR_PI = 16.0 * np.arctan(1.0)


def _g(x) -> str:
    """A double as std::ostream << double prints it (default precision 6)."""
    s = f"{float(x):g}"
    return s


def _vec(v) -> str:
    return "[" + ",".join(_g(c) for c in v) + "]"


def write_cell_counts(path, counts, n):
    """outputNumberOfParticlesPerCell (cell.h:1360-1395): the size, then i,j,k:count in
    row-major order, then the total."""
    n = tuple(int(v) for v in n)
    counts = np.asarray(counts, dtype=np.int64).reshape(-1)
    assert counts.size == n[0] * n[1] * n[2]
    lines = [f"[{n[0]},{n[1]},{n[2]}]"]
    i = 0
    for x in range(n[0]):
        for y in range(n[1]):
            for z in range(n[2]):
                lines.append(f"{x},{y},{z}:{int(counts[i])}")
                i += 1
    text = "\n".join(lines) + "\n" + f"\nTotal: {int(counts.sum())}\n"
    with open(path, "w") as f:
        f.write(text)


def read_cell_counts(path):
    """(n, counts in row-major order) of a per-cell count file."""
    with open(path) as f:
        rows = f.read().splitlines()
    n = tuple(int(v) for v in rows[0].strip("[]").split(","))
    counts = np.zeros(n[0] * n[1] * n[2], dtype=np.int64)
    for r in rows[1:]:
        if ":" not in r or r.startswith("Total"):
            continue
        key, val = r.split(":")
        x, y, z = (int(v) for v in key.split(","))
        counts[(x * n[1] + y) * n[2] + z] = int(val)
    return n, counts


def write_particle_positions(path, pos):
    """outputParticlePositions (cell.h:1397-1429): one "x y z" line per particle, cells in
    row-major order."""
    with open(path, "w") as f:
        for p in np.asarray(pos, dtype=np.float64).reshape(-1, 3):
            f.write(f"{_g(p[0])} {_g(p[1])} {_g(p[2])}\n")


def reference_field_energy(F, cell_width):
    """getFieldEnergy (field.h:443-459): sum over nodes of 0.5 vol |F|^2 / (4 pi)."""
    vol = 0.5 * float(np.prod(cell_width))
    F = np.asarray(F, dtype=np.float64).reshape(-1, 3)
    return float(np.sum(vol * np.sum(F * F, axis=1) / FOUR_PI))


class ConservedLog:
    """writeOutputHeader / writeOutputData (simulator.h:67-91): a tab-separated row of the
    cycle, total momentum, E and B field energies and total kinetic energy, every
    FieldOutputCycle cycles and at the last one."""

    HEADER = "Cycle \t Total Moment \t E energy \t B energy \t Total KE \n"

    def __init__(self, path, field_output_cycle=1):
        self.f = open(path, "w")
        self.cycle_every = int(field_output_cycle)
        self.f.write(self.HEADER)

    def due(self, cycle, num_steps):
        return self.cycle_every > 0 and (cycle % self.cycle_every == 0 or cycle + 1 == num_steps)

    def write(self, cycle, momentum, e_energy, b_energy, kinetic):
        self.f.write(f"{cycle}\t{_g(momentum)}\t{_g(e_energy)}\t{_g(b_energy)}\t{_g(kinetic)}\n")

    def write_from(self, cycle, sim, cell_width=(1.0, 1.0, 1.0)):
        """One row from a Simulation: particle sums on the GPU, field energies from node E and
        node B as getFieldEnergy computes them."""
        c = sim.conservation()
        f = sim.download_fields()
        self.write(cycle, c["momentum"], reference_field_energy(f["E"], cell_width),
                   reference_field_energy(f["Bn"], cell_width), c["kinetic_energy"])

    def close(self):
        self.f.close()


def write_density_frame(prefix, t, counts, n, cell_width):
    """The density frame of writeOutputData (simulator.h:93-113): file prefix + %06d of t with
    rows t,x,y,z,density (x, y, z the cell's lower corner relative to the origin)."""
    n = tuple(int(v) for v in n)
    counts = np.asarray(counts, dtype=np.int64).reshape(-1)
    path = f"{prefix}{int(t):06d}"
    with open(path, "w") as f:
        f.write("t,x,y,z,density\n")
        i = 0
        for x in range(n[0]):
            for y in range(n[1]):
                for z in range(n[2]):
                    f.write(f"{t},{_g(x * cell_width[0])},{_g(y * cell_width[1])},{_g(z * cell_width[2])},"
                            f"{int(counts[i])}\n")
                    i += 1
    return path


def write_field_grids(path, E, Bn, Bc, n, Bext=None):
    """outputFieldGrids (field.h:464-530): the node grid as i,j,k:E|B|Bext, a blank line, then
    the centre grid as i,j,k:Bc (both written to the same file, as the reference does)."""
    n = tuple(int(v) for v in n)
    E = np.asarray(E, dtype=np.float64).reshape(-1, 3)
    Bn = np.asarray(Bn, dtype=np.float64).reshape(-1, 3)
    Bc = np.asarray(Bc, dtype=np.float64).reshape(-1, 3)
    Bext = np.zeros_like(E) if Bext is None else np.asarray(Bext, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        for grid in (0, 1):
            f.write(f"[{n[0]},{n[1]},{n[2]}]\n")
            i = 0
            for x in range(n[0]):
                for y in range(n[1]):
                    for z in range(n[2]):
                        if grid == 0:
                            f.write(f"{x},{y},{z}:{_vec(E[i])}|{_vec(Bn[i])}|{_vec(Bext[i])}\n")
                        else:
                            f.write(f"{x},{y},{z}:{_vec(Bc[i])}\n")
                        i += 1
            f.write("\n")
