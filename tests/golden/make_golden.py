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
    with open(os.path.join(REF, "outputs/test.out")) as f:
        data = f.read()
    with open(os.path.join(HERE, "test_out_counts.txt"), "w") as f:
        f.write(data)


if __name__ == "__main__":
    main()
