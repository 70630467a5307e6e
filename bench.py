#!/usr/bin/env python3
"""Benchmark of the MI355X particle-in-cell step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 128] [--ppc 32]

Workload (default): 128^3 periodic cells x 32 particles per cell (67,108,864 particles), two
species as in the two-species configuration (electrons q/m=-1, vth 0.05; ions q/m=1/64,
vth 0.006; 16 + 16 per cell), B=(0,0.2,1), E ~ N(0, 1e-3^2) per node component, dt=0.05,
plain FP32 leapfrog.  Synthetic data generated on the GPU.

One step = leapfrog E and B passes + interpC2N (with the FP64 cross-tile J sum), the fused
gather + Boris + rebin + deposit pass over every particle, and the cell-count prefix sums.
Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle updates/s (push+deposit) & % HBM peak; surrogate speedup at ≤1e-3 err"
HBM_PEAK_GBS = 8000.0
FOUR_PI = 16.0 * np.arctan(1.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_particle(ppc: float) -> float:
    # FP32 offsets xyz + velocities xyz read and written once (48 B), plus per-node E,B
    # gather (24 B) and J write (12 B) amortised over the particles of a cell
    return 48.0 + 36.0 / ppc


def load_traffic(cells: int, ppc: int):
    """HBM bytes per push launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), recorded in profiles/ for this workload."""
    path = os.path.join(ROOT, "profiles", f"push_kernel_{cells}c_{ppc}ppc_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def make_sim(G, cells, ppc, seed, flags=0, tile=(0, 0, 0)):
    half = ppc // 2
    q = 1.0 / (half * FOUR_PI)
    sim = G.Simulation(cells, dt=0.05, c=1.0, species=((-q, -1.0), (q, 1.0 / 64)), flags=flags, tile=tile)
    sim.init_particles(0, half, (0.05,) * 3, seed=seed)
    sim.init_particles(1, ppc - half, (0.006,) * 3, seed=seed)
    rng = np.random.default_rng(seed)
    n3 = cells ** 3
    E = rng.normal(0.0, 1e-3, size=(n3, 3))
    Bc = np.tile([0.0, 0.2, 1.0], (n3, 1))
    sim.upload_fields(E, Bc)
    return sim


def cpu_baseline(cells: int, ppc: int, budget_s: float = 12.0):
    """The FP64 oracle (a CPU restatement of the reference step) on this host's cores, on a
    bounded sample of the same workload: a 64^3 box (1/8 of the cells) with the same
    particles per cell and physics, timed over as many steps as fit in the budget."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    n = min(cells, 64)
    half = ppc // 2
    q = 1.0 / (half * FOUR_PI)
    u = O.Universe(n, dt=0.05)
    u.init_species(0, 1, half, -q, -1.0, (0.05,) * 3)
    u.init_species(1, 1, ppc - half, q, 1.0 / 64, (0.006,) * 3)
    rng = np.random.default_rng(1)
    u.set_fields(rng.normal(0.0, 1e-3, size=(n ** 3, 3)), np.tile([0.0, 0.2, 1.0], (n ** 3, 1)))
    npart = u.num_particles()

    def timed(nthreads, budget, max_steps):
        steps, t0 = 0, time.perf_counter()
        while True:
            u.step(1, nthreads)
            steps += 1
            el = time.perf_counter() - t0
            if el > budget or steps >= max_steps:
                return steps, el

    u.step(1, threads)  # warm-up
    steps, el = timed(threads, budget_s, 40)
    steps1, el1 = timed(1, budget_s / 2, 5)  # the same step on one core
    u.close()
    return {"value": npart * steps / el, "unit": "particle updates/s", "cores": threads, "kind": "port",
            "value_1core": npart * steps1 / el1,
            "sample": f"FP64 oracle step (deposit, leapfrog, gather+Boris, rebin), {n}^3 cells x {ppc} ppc "
                      f"({npart} particles), {steps} steps in {el:.1f} s on {threads} threads; "
                      f"{steps1} steps in {el1:.1f} s on 1 thread"}


def copy_peak_gbs(G):
    """Measured device-to-device copy bandwidth (read + write bytes / time), via torch."""
    try:
        import torch
        if not torch.cuda.is_available():
            log("copy peak unavailable: torch sees no GPU")
            return None
        n = 1 << 28  # 1 GiB of fp32
        a = torch.empty(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(10):
            b.copy_(a)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / 10
        return 2 * n * 4 / (ms * 1e-3) / 1e9
    except Exception as ex:  # pragma: no cover
        log("copy peak unavailable:", ex)
        return None


OPERATOR_FILE = os.path.join(ROOT, "allscale_ipic3d_amd", "data", "operator_64.npz")


def surrogate_leg(G, cells, ppc, phases=10, seed=21):
    """Field phase of the learned operator against 8 leapfrog substeps on the same state.

    Times one field phase each way with HIP events on the library's stream (the operator's
    phase includes the J sum, pack, network, head, the div B / energy checks with their host
    read-back, and interpC2N), then audits every operator application against leapfrog for
    the error |F_op - F_lf| / |F_lf| (total field) and the increment-relative error."""
    from allscale_ipic3d_amd import field_operator as FO
    trained = os.path.exists(OPERATOR_FILE)
    if trained:
        from allscale_ipic3d_amd.train_operator import load
        params, scales, meta = load(OPERATOR_FILE)
    else:
        params, scales, meta = FO.init_params(0), {"in_E": 1e-3, "in_B": 1e-3, "in_J": 1e-3, "out_E": 1e-5,
                                                   "out_P": 5e-5}, {}
    sim = FO.config2_simulation(cells, ppc, seed, field_substeps=8, flags=G.FLAG_PHASE_TIMING)
    sim.step(2)
    lf = G.StepStats()
    for _ in range(phases):
        sim.field_step(use_operator=False, stats=lf)
        sim.particle_step()
    sim.set_field_operator(params, scales, divb_tol=1e-4, energy_tol=1e-3)
    op = G.StepStats()
    for _ in range(phases):
        sim.field_step(use_operator=True, stats=op)
        sim.particle_step()
    sim.set_field_operator(params, scales, divb_tol=1e30, energy_tol=1e30, audit_every=1, audit_tol=1e30)
    au = G.StepStats()
    for _ in range(phases):
        sim.field_step(use_operator=True, stats=au)
        sim.particle_step()
    sim.close()
    t_lf, t_op = lf.ms_field / phases, op.ms_field / phases
    flops = FO.FLOP_PER_CELL * cells ** 3
    return {
        "workload": f"{cells}^3 cells x {ppc} ppc two-species field phase, 8 leapfrog substeps (J fixed) vs one "
                    "operator application", "trained": trained, "training": meta,
        "t_leapfrog_ms": t_lf, "t_operator_ms": t_op, "speedup": t_lf / t_op if t_op > 0 else None,
        "operator_steps": int(op.operator_steps), "fallback_steps": int(op.fallback_steps),
        "max_divb_residual": op.max_divb_residual, "max_energy_residual": op.max_energy_residual,
        "rel_error_total": au.max_audit_error, "rel_error_increment": au.max_audit_error_increment,
        "flop_per_inference": flops, "operator_phase_tflops": flops / (t_op * 1e-3) / 1e12 if t_op > 0 else None,
    }


def other_configs(G):
    """Per-step times of the other configurations (parity-test cases, not the metric): the
    two-species 64^3 x 64 ppc box, the uneven-load 128^3 dipole configuration (30-299 ppc,
    74.9M particles, sink) and the reference main driver's inputs/test.inp and tiny.inp cases
    (dipole mover, reflection, sink, the reference's random streams)."""
    res = {}
    sim = make_sim(G, 64, 64, 5)
    sim.step(3)
    st = sim.step(20)
    t = (st["first_step_s"] + st["remaining_steps_s"]) / 20
    res["two_species_64c_64ppc"] = {"particles": 64 ** 3 * 64, "ms_per_step": t * 1e3,
                                     "updates_per_s": 64 ** 3 * 64 / t}
    sim.close()

    n = 128
    c = n / 2.0
    x = np.arange(n) + 0.5
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    d = np.stack([X - c, Y - c, Z - c], axis=-1).reshape(-1, 3)
    r2 = np.sum(d * d, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = 1.0 / r2 ** 2.5
    B = np.stack([3 * d[:, 0] * d[:, 2] * fac, 3 * d[:, 1] * d[:, 2] * fac,
                  (2 * d[:, 2] ** 2 - d[:, 0] ** 2 - d[:, 1] ** 2) * fac], axis=1)
    B[r2 <= 16.0] = 0.0
    ppc = np.rint(30 * (1 + 9 * np.exp(-r2 / 400.0))).astype(np.int32)
    q = -1.0 / (30 * FOUR_PI)
    sim = G.Simulation(n, dt=0.1, species=((q, -1.0),), sink_center=(c, c, c), sink_radius=4.0)
    sim.init_particles(0, ppc, (0.05,) * 3, drift=(0.1, 0.0, 0.0), seed=3)
    sim.upload_fields(np.zeros((n ** 3, 3)), B)
    sim.step(2)
    st = sim.step(5)
    t = (st["first_step_s"] + st["remaining_steps_s"]) / 5
    res["uneven_dipole_128c"] = {"particles": int(ppc.sum()), "ppc_range": [int(ppc.min()), int(ppc.max())],
                                 "ms_per_step": t * 1e3, "updates_per_s": float(ppc.sum()) / t}
    sim.close()

    from allscale_ipic3d_amd import driver as D
    for name in ("test", "tiny"):
        inp = os.path.join(ROOT, "tests", "golden", f"{name}_case.inp")
        if not os.path.exists(inp):
            continue
        case = D.driver_case(D.Parameters(inp), name)
        sim = D.make_simulation(case)
        sim.init_reference(0, D.case_distribution(case))
        sim.step(2)
        st = sim.step(20)
        t = (st["first_step_s"] + st["remaining_steps_s"]) / 20
        res[f"reference_driver_{name}"] = {"cells": list(case.n), "particles": case.num_particles,
                                           "ms_per_step": t * 1e3, "updates_per_s": case.num_particles / t}
        sim.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tile", type=int, nargs=3, default=(0, 0, 0),
                    help="cells per tile along x y z (0: the library's default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy-peak", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned field step leg")
    ap.add_argument("--surrogate-cells", type=int, default=64)
    ap.add_argument("--no-other-configs", action="store_true", help="skip the per-step times of the other configurations")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 10 times (rocprof counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import allscale_ipic3d_amd as G

    if args.calibrate_copy:
        import torch
        n = 1 << 28
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 10}))
        return

    copy_peak = None if args.no_copy_peak else copy_peak_gbs(G)
    info = G.device_info()
    log("device:", info)
    cells, ppc = args.cells, args.ppc
    npart = cells ** 3 * ppc

    log(f"init {cells}^3 x {ppc} ppc = {npart} particles")
    sim = make_sim(G, cells, ppc, args.seed, tile=tuple(args.tile))
    if args.warmup:
        sim.step(args.warmup)
    log(f"warmup {args.warmup} steps done")
    st = sim.step(args.steps)  # the call synchronises the stream before and after
    t_total = st["first_step_s"] + st["remaining_steps_s"]
    value = npart * args.steps / t_total
    log(f"timed {args.steps} steps: {t_total * 1e3:.2f} ms, {value:.4e} particle updates/s")
    live = sim.num_particles()
    sim.close()

    # the dominant kernel's launch duration, measured with HIP events on the library's stream
    simt = make_sim(G, cells, ppc, args.seed, flags=G.FLAG_PHASE_TIMING, tile=tuple(args.tile))
    simt.step(max(2, args.warmup))
    nt = min(args.steps, 20)
    stt = simt.step(nt)
    simt.close()
    push_ms = stt["ms_push"] / nt
    field_ms = stt["ms_field"] / nt
    scan_ms = stt["ms_scan"] / nt
    Bp = algorithmic_bytes_per_particle(ppc)
    alg_bytes = Bp * npart
    achieved = alg_bytes / (push_ms * 1e-3) / 1e9
    log(f"push kernel {push_ms:.3f} ms/launch -> {achieved:.1f} GB/s algorithmic "
        f"({achieved / HBM_PEAK_GBS * 100:.1f}% of {HBM_PEAK_GBS:.0f} GB/s); field {field_ms:.3f} ms, scan {scan_ms:.3f} ms")

    sur = None
    if not args.no_surrogate:
        sur = surrogate_leg(G, args.surrogate_cells, 64)
        log("surrogate:", {k: v for k, v in sur.items() if k != "training"})
    others = None
    if not args.no_other_configs:
        others = other_configs(G)
        log("other configurations:", others)
    cpu = None if args.no_cpu_baseline else cpu_baseline(cells, ppc)
    if cpu:
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_total * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain PIC step, {cells}^3 periodic cells x {ppc} particles/cell ({npart} particles), "
                        "two species (q/m -1 and 1/64, vth 0.05 and 0.006), B=(0,0.2,1), E~N(0,1e-6), dt=0.05, "
                        "FP32 mover/deposit with FP64 cross-tile J sum",
            "cells": [cells] * 3,
            "particles_per_cell": ppc,
            "particles": npart,
            "live_particles_after": live,
            "parallelism": "single GPU",
            "device": info["name"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_particles<PUSH> (gather + Boris + rebin + deposit)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(cells, ppc),
            "algorithmic_bytes_per_particle": Bp,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_ms": push_ms,
            "measured_copy_peak_gbs": copy_peak,
            "step_breakdown_ms": {"field": field_ms, "push": push_ms, "scan": scan_ms},
        },
        "cpu_baseline": cpu,
        "surrogate": sur,
        "other_configs": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
