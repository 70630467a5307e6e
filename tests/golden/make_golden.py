"""Extracts the golden vectors used by the oracle tests from the reference's own data files.

Run in the build container only (it reads /root/reference, which GPU hosts do not have):
    python tests/golden/make_golden.py
Outputs (committed fixtures, data only):
  larmor_radius_rows.txt  -- every 50th row and the last 5 rows of results/larmor_radius.txt
                             (FP64 Boris gyration, 1000 steps; row 999 is asserted by
                             code/app/test/simulation_test.cc:255-261)
  exb_drift_rows.txt      -- every 100th row and the last 5 rows of results/exb_drift.txt
                             (E x B drift, 5000 steps; simulation_test.cc:399-406)
  test_out_counts.txt     -- outputs/test.out: per-cell particle counts after the dipole-mode
                             run of inputs/test.inp (scripts/verify_output.sh.in:48 oracle)
  tiny_out_counts.txt     -- outputs/tiny.out, the same for inputs/tiny.inp
  test_conserved.txt      -- outputs/testConservedQuantities.out: per-cycle momentum, E and B
  tiny_conserved.txt         field energies and kinetic energy of those runs (simulator.h:73-91);
                             tiny_conserved.txt from outputs/tinyConservedQuantities.out
  test_conserved_static.txt  -- outputs/testConservedQuantities.static.out and
  tiny_conserved_static.txt     tinyConservedQuantities.static.out: the same columns for runs whose
                             particles come from the parameter-file initCells (cell.h:699-776,
                             5^3 and 8^3 particles per cell, qom -25, uth 0.045, u0 0.02) in the
                             constant field of B0 and the drift (row 0 is the initial state)
  test_case.inp           -- the key = value lines of inputs/test.inp / tiny.inp that the
  tiny_case.inp              driver reads (comments and unused keys dropped)
Row format of the two trajectory files (as in the reference): step t x y z vx vy vz.
"""
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def subset(src, dst, every):
    with open(os.path.join(REF, src)) as f:
        rows = [ln for ln in f.read().splitlines() if ln.strip()]
    keep = [r for i, r in enumerate(rows) if i % every == 0 or i >= len(rows) - 5]
    with open(os.path.join(HERE, dst), "w") as f:
        f.write("\n".join(keep) + "\n")


def main():
    subset("results/larmor_radius.txt", "larmor_radius_rows.txt", 50)
    subset("results/exb_drift.txt", "exb_drift_rows.txt", 100)
    for src, dst in (("outputs/test.out", "test_out_counts.txt"), ("outputs/tiny.out", "tiny_out_counts.txt"),
                     ("outputs/testConservedQuantities.out", "test_conserved.txt"),
                     ("outputs/tinyConservedQuantities.out", "tiny_conserved.txt"),
                     ("outputs/testConservedQuantities.static.out", "test_conserved_static.txt"),
                     ("outputs/tinyConservedQuantities.static.out", "tiny_conserved_static.txt")):
        with open(os.path.join(REF, src)) as f:
            data = f.read()
        with open(os.path.join(HERE, dst), "w") as f:
            f.write(data)
    keys = ("dt", "ncycles", "Lx", "Ly", "Lz", "nxc", "nyc", "nzc", "B0x", "B0y", "B0z", "B1x", "B1y", "B1z",
            "npcelx", "npcely", "npcelz", "u0", "v0", "w0", "FieldOutputCycle", "ParticlesOutputCycle")
    for case in ("test", "tiny"):
        rows = []
        with open(os.path.join(REF, "inputs", case + ".inp")) as f:
            for ln in f:
                body = ln.split("#")[0]
                if "=" not in body:
                    continue
                k, v = (t.strip() for t in body.split("=", 1))
                if k in keys:
                    rows.append(f"{k} = {' '.join(v.split())}")
        with open(os.path.join(HERE, case + "_case.inp"), "w") as f:
            f.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main()
