"""The reference's output formats (allscale_ipic3d_amd/io.py): the per-cell count writer
reproduces the reference's own outputs/test.out byte for byte from its counts; the other
writers follow the layouts of simulator.h:67-115, cell.h:1397-1429 and field.h:464-530."""
import os

import numpy as np

from allscale_ipic3d_amd import io as IO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_out_counts.txt")


def test_cell_count_file_reproduces_reference_output(tmp_path):
    n, counts = IO.read_cell_counts(GOLD)
    assert n == (8, 8, 8) and counts.sum() == 256987
    out = tmp_path / "counts.out"
    IO.write_cell_counts(out, counts, n)
    assert out.read_bytes() == open(GOLD, "rb").read()


def test_conserved_log_layout(tmp_path):
    p = tmp_path / "log.txt"
    log = IO.ConservedLog(p, field_output_cycle=2)
    assert [c for c in range(5) if log.due(c, 5)] == [0, 2, 4]
    log.write(0, 1.5, 2.25e-7, 3.0, 1234567.0)
    log.close()
    rows = p.read_text().splitlines()
    assert rows[0] == "Cycle \t Total Moment \t E energy \t B energy \t Total KE "
    assert rows[1] == "0\t1.5\t2.25e-07\t3\t1.23457e+06"


def test_density_frame_and_positions(tmp_path):
    path = IO.write_density_frame(str(tmp_path / "dens"), 3, [1, 2, 3, 4, 5, 6, 7, 8], (2, 2, 2), (0.5, 0.5, 0.5))
    assert path.endswith("dens000003")
    rows = open(path).read().splitlines()
    assert rows[0] == "t,x,y,z,density" and rows[1] == "3,0,0,0,1" and rows[-1] == "3,0.5,0.5,0.5,8"
    q = tmp_path / "pos.txt"
    IO.write_particle_positions(q, [[1.0, 2.5, -3.0]])
    assert q.read_text() == "1 2.5 -3\n"


def test_reference_field_energy():
    # 0.5 * vol * |F|^2 / (4 pi) summed over nodes (field.h:443-459)
    F = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 3.0]])
    want = 0.5 * 2.0 * (9.0 + 9.0) / (16.0 * np.arctan(1.0))
    assert abs(IO.reference_field_energy(F, (1.0, 2.0, 1.0)) - want) < 1e-15


def test_field_grid_layout(tmp_path):
    p = tmp_path / "fields.txt"
    E = np.array([[1.0, 0.0, 0.5]])
    IO.write_field_grids(p, E, [[0, 0, 1]], [[0.25, 0, 0]], (1, 1, 1))
    assert p.read_text() == "[1,1,1]\n0,0,0:[1,0,0.5]|[0,0,1]|[0,0,0]\n\n[1,1,1]\n0,0,0:[0.25,0,0]\n\n"
