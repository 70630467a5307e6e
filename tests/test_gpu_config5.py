"""The full chosen form at scale as a run: 128^3 cells x 32 particles per cell (67,108,864), two
species (electrons q/m -1, vth 0.05; ions q/m 1/64, vth 0.006), B0 = (0, 0.2, 1), E ~ N(0, 1e-3^2),
dt = 0.05, 8 leapfrog substeps per field phase, from one seed, three ways:

  * plain FP32 GPU leapfrog (ipic_gpu_step),
  * the same with the learned field step (allscale_ipic3d_amd/data/operator_128.npz, the config-4
    architecture retrained at 128^3) in place of the 8 substeps, with its div B / energy fallback,
  * the FP64 oracle (leapfrog), started from the GPU's particles and fields.

Every IPIC_CONFIG5_EVERY steps (default 50) it prints one JSON line: the leapfrog run against
FP64 (fields in relative l2, per-cell particle sums, conservation sums, as in
test_gpu_fullsize.py), the evolving-field error |F - F_ref| / |F_ref - F0| (F0 the uniform
background, |F|^2 = |E|^2 + |Bc|^2) of the surrogate run against the leapfrog run and against
FP64, and the wall time per step of both GPU runs.  The leapfrog run must stay within 1e-5 of
FP64; the surrogate run's errors are reported, with a loose bound (25 % of the evolving field)
that only catches a broken operator: the operator's per-phase error (5e-4) accumulates, and the
500-step run measured 0.10 at its end.

It takes about 15 minutes (the FP64 oracle steps 67M particles on the host cores), so it runs
only when IPIC_CONFIG5_STEPS is set (500: the configuration's length)."""
import json
import os

import numpy as np
import pytest

import oracle.compare as OC

pytestmark = pytest.mark.gpu

STEPS = int(os.environ.get("IPIC_CONFIG5_STEPS", "0"))
EVERY = int(os.environ.get("IPIC_CONFIG5_EVERY", "50"))
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OPERATOR = os.path.join(ROOT, "allscale_ipic3d_amd", "data", "operator_128.npz")
TOL = 1e-5
B0 = np.array([0.0, 0.2, 1.0])


def _evolving(got, ref):
    num = np.sum((got["E"] - ref["E"]) ** 2) + np.sum((got["Bc"] - ref["Bc"]) ** 2)
    den = np.sum(ref["E"] ** 2) + np.sum((ref["Bc"] - B0) ** 2)
    return float(np.sqrt(num / den))


@pytest.mark.skipif(STEPS <= 0, reason="a 15-minute run: set IPIC_CONFIG5_STEPS (500 for the full configuration)")
@pytest.mark.timeout(3600)
def test_config5_surrogate_and_leapfrog_against_fp64():
    from allscale_ipic3d_amd import field_operator as FO
    from allscale_ipic3d_amd.train_operator import load
    n, ppc, seed = 128, 32, 1
    params, scales, meta = load(OPERATOR)
    assert meta["n"] == [n, n, n] and meta["ppc"] == ppc
    lf = FO.config2_simulation(n, ppc, seed, field_substeps=8)
    su = FO.config2_simulation(n, ppc, seed, field_substeps=8)
    su.set_field_operator(params, scales, divb_tol=1e-4, energy_tol=1e-3)
    half = ppc // 2
    q = 1.0 / (half * 16.0 * np.arctan(1.0))
    species = ((-q, -1.0), (q, 1.0 / 64))
    rng = np.random.default_rng(seed)          # config2_simulation's initial field
    E = rng.normal(0.0, 1e-3, size=(n ** 3, 3))
    Bc = np.tile(B0, (n ** 3, 1))
    u = OC.oracle_from_gpu(lf, species, E, Bc, dt=0.05, c=1.0, field_substeps=8)
    npart = n ** 3 * ppc
    done = 0
    t_lf = t_su = 0.0
    ops = fallbacks = 0
    while done < STEPS:
        k = min(EVERY, STEPS - done)
        st = lf.step(k)
        t_lf += st["first_step_s"] + st["remaining_steps_s"]
        st = su.step(k)
        t_su += st["first_step_s"] + st["remaining_steps_s"]
        ops += int(st["operator_steps"])
        fallbacks += int(st["fallback_steps"])
        u.step(k)
        done += k
        res = OC.compare(u, lf, species)
        f_lf, f_su, f_ref = lf.download_fields(), su.download_fields(), u.fields()
        c_lf, c_su = lf.conservation(), su.conservation()
        row = {"step": done, "leapfrog_vs_fp64": res,
               "leapfrog_vs_fp64_evolving": _evolving(f_lf, f_ref),
               "surrogate_vs_leapfrog_evolving": _evolving(f_su, f_lf),
               "surrogate_vs_fp64_evolving": _evolving(f_su, f_ref),
               "surrogate_kinetic_energy_vs_leapfrog": abs(c_su["kinetic_energy"] / c_lf["kinetic_energy"] - 1),
               "surrogate_operator_steps": ops, "surrogate_fallback_steps": fallbacks,
               "ms_per_step_leapfrog": 1e3 * t_lf / done, "ms_per_step_surrogate": 1e3 * t_su / done}
        print(json.dumps(row), flush=True)
        for key, e in res["fields"].items():
            assert e <= TOL, f"step {done}: leapfrog {key} rel error {e:.3e}"
        for s, p in enumerate(res["particles"]):
            assert p["particles"] == p["particles_ref"]
            assert p["offset_max"] <= TOL and p["velocity_max"] <= TOL, f"step {done} species {s}: {p}"
            assert p["flip_particles"] <= 1e-4 * npart
        for key in ("kinetic_energy", "momentum", "field_energy_E", "field_energy_B"):
            assert res["conservation"][key] <= TOL, f"step {done}: {key}"
        assert row["surrogate_vs_leapfrog_evolving"] < 0.25
        del f_lf, f_su, f_ref
    assert ops > 0
    lf.close()
    su.close()
    u.close()
