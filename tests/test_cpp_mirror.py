"""The C++ mirror of the reference's step API (include/ipic3d_gpu/ipic3d.hpp), driven by
tests/cpp/ipic3d_mirror_test.cpp: without a GPU its Universe reports the HIP error code;
on the GPU the same case run through C++ and through the Python binding gives the same
per-cell counts and kinetic energy, and the field-operator policy guard fires."""
import os
import struct
import subprocess

import numpy as np
import pytest

import allscale_ipic3d_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "ipic3d_mirror_test")
need_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="C++ mirror test not built")


def _has_gpu():
    try:
        G.device_info()
        return True
    except Exception:
        return False


@need_exe
def test_cpp_mirror_reports_missing_device():
    if _has_gpu():
        pytest.skip("a GPU is present")
    r = subprocess.run([EXE, "--expect-no-device"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert '"error": -2' in r.stdout


def _case(n=8, ppc=8, seed=4):
    rng = np.random.default_rng(seed)
    nc = n ** 3
    cells = np.repeat(np.arange(nc), ppc)
    x = cells // (n * n)
    y = (cells // n) % n
    z = cells % n
    pos = np.stack([x, y, z], axis=1) + rng.uniform(0.0, 1.0, size=(len(cells), 3))
    vel = rng.normal(0.0, 0.05, size=(len(cells), 3))
    species = [(-0.01, -1.0), (0.01, 1.0 / 64)]
    kind = rng.integers(0, 2, size=len(cells))
    E = rng.normal(0.0, 1e-3, size=(nc, 3))
    Bc = np.tile([0.0, 0.2, 1.0], (nc, 1))
    return n, pos, vel, species, kind, E, Bc


@pytest.mark.gpu
@need_exe
def test_cpp_mirror_matches_python_binding(tmp_path):
    n, pos, vel, species, kind, E, Bc = _case()
    steps, dt = 5, 0.1
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<qqqd", n, steps, len(pos), dt))
        rec = np.zeros((len(pos), 8))
        rec[:, 0:3], rec[:, 3:6] = pos, vel
        rec[:, 6] = [species[k][0] for k in kind]
        rec[:, 7] = [species[k][1] for k in kind]
        f.write(rec.astype("<f8").tobytes())
        f.write(E.astype("<f8").tobytes())
        f.write(Bc.astype("<f8").tobytes())
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = out.read_bytes()
    nc = n ** 3
    count = struct.unpack_from("<q", raw, 0)[0]
    cells = np.frombuffer(raw, dtype="<i8", count=nc, offset=8)
    ke, first, rest, guard = struct.unpack_from("<dddq", raw, 8 + 8 * nc)
    assert guard == 1
    assert count == len(pos)
    assert first >= 0 and rest >= 0

    # the same case through the Python binding (species in the C++ program's sorted order)
    order = sorted(set(species))
    sim = G.Simulation(n, dt=dt, species=order)
    for s, sp in enumerate(order):
        sel = np.array([species[k] == sp for k in kind])
        sim.upload_particles(s, pos[sel], vel[sel])
    sim.upload_fields(E, Bc)
    sim.step(steps)
    np.testing.assert_array_equal(cells, sim.cell_counts())
    ke_py = sim.conservation()["kinetic_energy"]
    assert abs(ke - ke_py) <= 1e-12 * abs(ke_py)
    # the reference's output formats from C++ (io.hpp) and Python (io.py) agree
    from allscale_ipic3d_amd import io as IO
    IO.write_cell_counts(tmp_path / "py.counts.out", sim.cell_counts(), (n, n, n))
    assert (tmp_path / "out.bin.counts.out").read_bytes() == (tmp_path / "py.counts.out").read_bytes()
    rows = (tmp_path / "out.bin.log").read_text().splitlines()
    assert rows[0].startswith("Cycle \t Total Moment") and len(rows) == 3
    assert rows[2].split("\t")[0] == str(steps - 1)
    sim.close()


@pytest.mark.gpu
@need_exe
def test_cpp_mirror_driver_case(tmp_path):
    """createUniverseFromDistribution + simulateSteps through the C++ mirror on the main driver's
    test.inp set-up gives the per-cell counts of the Python driver on the same case."""
    from allscale_ipic3d_amd import driver as D, io as IO
    prefix = str(tmp_path / "cpp_test")
    r = subprocess.run([EXE, "--driver-case", prefix], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    n, cells = IO.read_cell_counts(prefix + ".out")
    assert n == (8, 8, 8)
    res = D.run(os.path.join(ROOT, "tests", "golden", "test_case.inp"), write=False)
    np.testing.assert_array_equal(cells, res["counts"])


@pytest.mark.gpu
@need_exe
@pytest.mark.parametrize("case", ["test", "tiny"])
def test_cpp_mirror_params_case(case):
    """createUniverseFromParams (parameter-file initCells + initFields) through the C++ mirror
    gives row 0 of outputs/{test,tiny}ConservedQuantities.static.out: 8^3 cells x 5^3 and 16^3 x
    8^3 lattice particles."""
    import json
    n, ppc, gold = {"test": (8, 5, "test_conserved_static.txt"), "tiny": (16, 8, "tiny_conserved_static.txt")}[case]
    r = subprocess.run([EXE, "--params-case", str(n), str(ppc)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = json.loads(r.stdout.strip().splitlines()[-1])
    with open(os.path.join(ROOT, "tests", "golden", gold)) as f:
        mom, e_en, b_en, ke = (float(v) for v in f.read().splitlines()[1].split()[1:])
    assert got["particles"] == n ** 3 * ppc ** 3
    assert abs(got["momentum"] / mom - 1) < 1e-5 and abs(got["kinetic_energy"] / ke - 1) < 1e-5
    assert abs(got["E_energy"] / e_en - 1) < 5e-6
    assert got["B_energy"] <= b_en and abs(got["B_energy"] / b_en - 1) < 1e-3
