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
BF16_PEAK_FLOPS = 2.5e15  # MI355X dense BF16 MFMA
FOUR_PI = 16.0 * np.arctan(1.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_particle(ppc: float) -> float:
    # FP32 offsets xyz + velocities xyz read and written once (48 B), plus per-node E,B
    # gather (24 B) and J write (12 B) amortised over the particles of a cell
    return 48.0 + 36.0 / ppc


PUSH_SOURCE = os.path.join(ROOT, "allscale_ipic3d_amd", "csrc", "ipic_gpu.hip")


PUSH_SECTION_END = b"// field kernels (N-periodic, row-major float4)"


def source_sha256(text: bytes = None) -> str:
    """Hash of the push kernel's device code: csrc/ipic_gpu.hip from its start up to the field
    kernels' banner (constants, helpers, k_particles).  Host-side edits further down leave it as
    is; they are not covered, the launch set-up included."""
    import hashlib
    if text is None:
        with open(PUSH_SOURCE, "rb") as f:
            text = f.read()
    end = text.find(PUSH_SECTION_END)
    if end < 0:
        raise RuntimeError("push kernel section banner not found in " + PUSH_SOURCE)
    return hashlib.sha256(text[:end]).hexdigest()


def load_traffic(cells: int, ppc: int):
    """HBM bytes per push launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration), recorded in profiles/ for this workload.  The record carries
    the hash of the kernel source it was measured on; a record of another source is stale and
    gives no figure.  Returns (bytes or None, status)."""
    path = os.path.join(ROOT, "profiles", f"push_kernel_{cells}c_{ppc}ppc_pmc.json")
    if not os.path.exists(path):
        return None, "no record"
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, "unreadable record"
    if d.get("source_sha256") != source_sha256():
        return None, "stale: measured on another version of the push kernel (csrc/ipic_gpu.hip)"
    return float(d["hbm_bytes_per_launch"]), "measured on this source"


def pmc_summary(fetch_csv: str, write_csv: str, fetch_scale: float, write_scale: float, out: str, command: str):
    """Summarises two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, WRITE_SIZE, KiB per
    dispatch) of the push kernel into the record load_traffic reads, stamped with the source hash."""
    import csv

    def per_dispatch(path, counter):
        vals = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if "k_particles<0" in row["Kernel_Name"] and row["Counter_Name"] == counter:
                    key = row.get("Dispatch_Id") or row.get("Correlation_Id")
                    vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
        v = list(vals.values())
        return {"dispatches": len(v), "avg_kib": sum(v) / len(v), "min_kib": min(v), "max_kib": max(v)}

    fe = per_dispatch(fetch_csv, "FETCH_SIZE")
    wr = per_dispatch(write_csv, "WRITE_SIZE")
    rd, wb = fe["avg_kib"] * 1024 * fetch_scale, wr["avg_kib"] * 1024 * write_scale
    rec = {"command": command, "kernel": "k_particles<0 (push)", "source_sha256": source_sha256(),
           "FETCH_SIZE_kib_per_dispatch": fe, "WRITE_SIZE_kib_per_dispatch": wr,
           "fetch_scale": fetch_scale, "write_scale": write_scale,
           "note": "scales from a 1 GiB device copy (python bench.py --calibrate-copy under the same counters)",
           "hbm_read_bytes_per_launch": rd, "hbm_write_bytes_per_launch": wb, "hbm_bytes_per_launch": rd + wb,
           "source_sha256_covers": "csrc/ipic_gpu.hip up to the field kernels banner (the push kernel device code)"}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))


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


def host_cpus():
    """CPUs this process may use (its affinity mask, capped by OMP_NUM_THREADS when set, which the
    GPU box sets to the job's CPU share), the machine's total and its model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        avail = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:  # pragma: no cover
        pass
    return {"threads": min(avail, omp) if omp > 0 else avail, "affinity_cpus": avail,
            "machine_cpus": os.cpu_count(), "omp_num_threads": omp or None, "model": model}


def cpu_leg(G, cells: int, ppc: int, seed: int, fp64_steps: int = 3, budget_s: float = 10.0):
    """The FP64 oracle (a CPU restatement of the reference step) on this host's cores.

    Full size: the oracle takes exactly the particles and fields of a GPU run of the benchmark
    workload (same seed; FP32 values widened to FP64); both advance `fp64_steps` steps.  The
    oracle's steps are timed on all the job's cores (the CPU baseline at the workload's size) and
    the two states are compared (fields, per-cell particle sums, conservation sums: the
    FP32-vs-FP64 error of the measured workload).  A 64^3 sample of the same physics is also timed
    over more steps, and on one core."""
    import oracle as O
    import oracle.compare as OC
    cpu = host_cpus()
    threads = cpu["threads"]
    half = ppc // 2
    q = 1.0 / (half * FOUR_PI)
    species = ((-q, -1.0), (q, 1.0 / 64))
    sim = make_sim(G, cells, ppc, seed)
    rng = np.random.default_rng(seed)
    n3 = cells ** 3
    E = rng.normal(0.0, 1e-3, size=(n3, 3))
    Bc = np.tile([0.0, 0.2, 1.0], (n3, 1))
    t0 = time.perf_counter()
    u = OC.oracle_from_gpu(sim, species, E, Bc, dt=0.05, c=1.0, field_substeps=1)
    log(f"oracle holds the GPU state ({u.num_particles()} particles) after {time.perf_counter() - t0:.1f} s")
    npart = u.num_particles()
    t0 = time.perf_counter()
    u.step(fp64_steps, threads)
    el_full = time.perf_counter() - t0
    sim.step(fp64_steps)
    res = OC.compare(u, sim, species)
    u.close()
    sim.close()
    worst = max([*res["fields"].values(), *[p[k] for p in res["particles"] for k in ("offset_max", "velocity_max")],
                 *[v for k, v in res["conservation"].items() if k not in ("particles", "charge")]])
    fp64 = {"workload": "the benchmark workload (same seed and inputs)", "steps": fp64_steps,
            "tolerance": 1e-5, "max_rel_error": worst, **res,
            "definition": "fields: relative l2 error of E, Bc, Bn, J; particles per species: offset_max = max over "
                          "cells of |sum of offsets GPU - FP64| / count (cells), velocity_max = max over cells of "
                          "|sum of v GPU - FP64| / sum of |v| (cells whose counts agree; flip_* = cells/particles "
                          "on the other side of a face); conservation: relative error of the FP64 sums"}
    log("fp64 error:", json.dumps(fp64))

    # 64^3 sample of the same physics: more steps, and one core
    n = min(cells, 64)
    us = O.Universe(n, dt=0.05)
    us.init_species(0, 1, half, -q, -1.0, (0.05,) * 3)
    us.init_species(1, 1, ppc - half, q, 1.0 / 64, (0.006,) * 3)
    rng = np.random.default_rng(1)
    us.set_fields(rng.normal(0.0, 1e-3, size=(n ** 3, 3)), np.tile([0.0, 0.2, 1.0], (n ** 3, 1)))
    ns = us.num_particles()

    def timed(nthreads, budget, max_steps):
        steps, t0 = 0, time.perf_counter()
        while True:
            us.step(1, nthreads)
            steps += 1
            el = time.perf_counter() - t0
            if el > budget or steps >= max_steps:
                return steps, el

    us.step(1, threads)  # warm-up
    steps, el = timed(threads, budget_s, 40)
    steps1, el1 = timed(1, budget_s / 2, 5)
    us.close()
    base = {"value": npart * fp64_steps / el_full, "unit": "particle updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle step (deposit, leapfrog, gather+Boris, rebin) on the full workload, "
                      f"{cells}^3 cells x {ppc} ppc ({npart} particles), {fp64_steps} steps in {el_full:.1f} s on "
                      f"{threads} threads",
            "host": cpu,
            "sample_64c": {"value": ns * steps / el, "cores": threads,
                           "sample": f"{n}^3 cells x {ppc} ppc ({ns} particles), {steps} steps in {el:.1f} s"},
            "value_1core": ns * steps1 / el1,
            "sample_1core": f"{n}^3 cells x {ppc} ppc, {steps1} steps in {el1:.1f} s on 1 thread"}
    return base, fp64


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
    roll = rollout(G, FO, params, scales, cells, ppc, seed, phases)
    return {
        "workload": f"{cells}^3 cells x {ppc} ppc two-species field phase, 8 leapfrog substeps (J fixed) vs one "
                    "operator application", "trained": trained, "training": meta,
        "t_leapfrog_ms": t_lf, "t_operator_ms": t_op, "speedup": t_lf / t_op if t_op > 0 else None,
        # at the dense BF16 MFMA peak the inference alone takes flops / peak: the most any
        # operator of this size can gain over the measured leapfrog phase
        "speedup_roofline_bound": t_lf / (flops / BF16_PEAK_FLOPS * 1e3),
        "operator_steps": int(op.operator_steps), "fallback_steps": int(op.fallback_steps),
        "max_divb_residual": op.max_divb_residual, "max_energy_residual": op.max_energy_residual,
        "rel_error_evolving": au.max_audit_error_evolving, "rel_error_increment": au.max_audit_error_increment,
        "rel_error_total_with_background": au.max_audit_error,
        "rollout": roll,
        "error_definitions": "per phase, against leapfrog from the same state: evolving = |F_op - F_lf| / |F_lf - F0| "
                             "(F0 the uniform background), increment = |F_op - F_lf| / |F_lf - F_before|, "
                             "total_with_background = |F_op - F_lf| / |F_lf|; |F|^2 = |E|^2 + c^2 |Bc|^2",
        "flop_per_inference": flops, "operator_phase_tflops": flops / (t_op * 1e-3) / 1e12 if t_op > 0 else None,
    }


def rollout(G, FO, params, scales, cells, ppc, seed, phases):
    """Two runs from the same state, one with the operator (checks and fallback on) and one with
    leapfrog only, `phases` steps each; the field error of the operator run at the end relative
    to the leapfrog run's evolving field and to its change over the steps."""
    a = FO.config2_simulation(cells, ppc, seed, field_substeps=8)
    b = FO.config2_simulation(cells, ppc, seed, field_substeps=8)
    a.step(2)
    b.step(2)
    f0 = b.download_fields()
    a.set_field_operator(params, scales, divb_tol=1e-4, energy_tol=1e-3)
    st = G.StepStats()
    for _ in range(phases):
        a.field_step(use_operator=True, stats=st)
        a.particle_step()
        b.field_step(use_operator=False)
        b.particle_step()
    fa, fb = a.download_fields(), b.download_fields()
    a.close()
    b.close()
    bg = {"E": np.zeros(3), "Bc": np.asarray([0.0, 0.2, 1.0])}  # the uniform background of the run
    num = sum(float(np.sum((fa[k] - fb[k]) ** 2)) for k in ("E", "Bc"))
    evo = sum(float(np.sum((fb[k] - bg[k]) ** 2)) for k in ("E", "Bc"))
    inc = sum(float(np.sum((fb[k] - f0[k]) ** 2)) for k in ("E", "Bc"))
    return {"phases": phases, "operator_steps": int(st.operator_steps), "fallback_steps": int(st.fallback_steps),
            "rel_error_evolving": (num / evo) ** 0.5, "rel_error_change": (num / inc) ** 0.5 if inc > 0 else None}


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
    r2 = ((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2).reshape(-1)
    ppc = np.rint(30 * (1 + 9 * np.exp(-r2 / 400.0))).astype(np.int32)
    q = -1.0 / (30 * FOUR_PI)
    sim = G.Simulation(n, dt=0.1, species=((q, -1.0),), sink_center=(c, c, c), sink_radius=4.0)
    sim.init_particles(0, ppc, (0.05,) * 3, drift=(0.1, 0.0, 0.0), seed=3)
    sim.init_fields(G.dipole_field_init(1.0, (c, c, c), 4.0))
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
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 --pmc counter files of the push kernel into the traffic record")
    ap.add_argument("--pmc-scales", nargs=2, type=float, default=(2.0, 1.0), metavar=("FETCH", "WRITE"))
    ap.add_argument("--pmc-command", default="", help="the profiling command, recorded with the summary")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.pmc_scales[0], args.pmc_scales[1],
                    os.path.join(ROOT, "profiles", f"push_kernel_{args.cells}c_{args.ppc}ppc_pmc.json"),
                    args.pmc_command)
        return

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
    cpu, fp64 = (None, None) if args.no_cpu_baseline else cpu_leg(G, cells, ppc, args.seed)
    if cpu:
        log("cpu baseline:", cpu)

    traffic, traffic_status = load_traffic(cells, ppc)
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
            "traffic": traffic,
            "traffic_status": traffic_status,
            "algorithmic_bytes_per_particle": Bp,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_ms": push_ms,
            "measured_copy_peak_gbs": copy_peak,
            "step_breakdown_ms": {"field": field_ms, "push": push_ms, "scan": scan_ms},
        },
        "cpu_baseline": cpu,
        "fp64_error": fp64,
        "surrogate": sur,
        "other_configs": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
