"""Size-independent comparison of a GPU run with the FP64 oracle (test infrastructure only:
tests/ and bench.py's cpu_baseline leg use it as the checker).

The GPU keeps a different particle order inside a cell, so particles are compared through
per-cell sums: with the same particles in a cell, the per-cell sums of the cell-relative
offsets and velocities differ by at most the sum of the per-particle differences.  Cells
whose counts differ (a particle that the FP32 and FP64 runs place on different sides of a
face) are counted and left out of the sums.
"""
from __future__ import annotations

import numpy as np

from . import Universe, make_particles


def oracle_from_gpu(sim, species, E, Bc, **props):
    """An oracle universe holding exactly the GPU's current particles (FP32 values widened to
    FP64) and the given fields.  species: [(q, qom)] in the GPU's species order."""
    u = Universe(sim.n, **props)
    for s, (q, qom) in enumerate(species):
        p = sim.download_particles(s)
        u.add_particles(make_particles(p["pos"], p["vel"], q, qom))
        del p
    u.set_fields(E, Bc)
    return u


def _cell_sums(pos, vel, cell, n, w, origin):
    ncell = int(np.prod(n))
    coords = np.stack(np.unravel_index(cell, n), -1).astype(np.float64)
    off = (pos - np.asarray(origin)) / np.asarray(w) - coords
    cnt = np.bincount(cell, minlength=ncell)
    so = np.stack([np.bincount(cell, weights=off[:, k], minlength=ncell) for k in range(3)], -1)
    sv = np.stack([np.bincount(cell, weights=vel[:, k], minlength=ncell) for k in range(3)], -1)
    sa = np.bincount(cell, weights=np.linalg.norm(vel, axis=1), minlength=ncell)
    return cnt, so, sv, sa


def particle_errors(u, sim, species, w=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """Per species: cells whose counts differ ('flip_cells', and the particles they move,
    'flip_particles'), and over the other cells the largest per-particle mean offset difference
    max |sum_gpu(off) - sum_ref(off)| / count (cell units, 'offset_max') and the largest
    velocity-sum difference relative to the cell's sum of speeds
    max |sum_gpu(v) - sum_ref(v)| / sum_ref(|v|) ('velocity_max')."""
    n = tuple(sim.n)
    parts, cell = u.particles()
    out = []
    for s, (_, qom) in enumerate(species):
        m = parts["qom"] == qom
        rc, ro, rv, ra = _cell_sums(parts["pos"][m], parts["vel"][m], cell[m], n, w, origin)
        g = sim.download_particles(s)
        gc, go, gv, _ = _cell_sums(g["pos"], g["vel"], g["cell"], n, w, origin)
        del g
        same = (gc == rc) & (rc > 0)
        flips = int(np.count_nonzero(gc != rc))
        do = np.abs(go - ro)[same].max(1) / rc[same] if same.any() else np.zeros(1)
        dv = np.linalg.norm(gv - rv, axis=1)[same] / np.maximum(ra[same], 1e-300) if same.any() else np.zeros(1)
        out.append({"particles": int(gc.sum()), "particles_ref": int(rc.sum()), "flip_cells": flips,
                    "flip_particles": int(np.abs(gc - rc).sum()) // 2,
                    "offset_max": float(do.max()), "velocity_max": float(dv.max())})
    del parts, cell
    return out


def rel_l2(got, ref):
    ref = np.asarray(ref, dtype=np.float64)
    den = float(np.linalg.norm(ref))
    return float(np.linalg.norm(np.asarray(got, dtype=np.float64) - ref) / (den if den > 0 else 1.0))


def field_errors(u, sim, names=("E", "Bc", "Bn", "J")):
    ref = u.fields()
    got = sim.download_fields()
    return {k: rel_l2(got[k], ref[k]) for k in names}


def conservation_errors(u, sim, species):
    """Relative errors of the FP64 sums; the charge relative to the summed magnitude of the
    species' charges (the total of a neutral plasma is zero up to rounding)."""
    c = sim.conservation()
    wE, wB = u.field_energy()
    ref = {"kinetic_energy": u.kinetic_energy(), "momentum": u.momentum(), "field_energy_E": wE,
           "field_energy_B": wB}
    out = {k: abs(c[k] - v) / abs(v) if v != 0 else abs(c[k]) for k, v in ref.items()}
    qabs = sum(abs(q) * sim.num_particles(s) for s, (q, _) in enumerate(species))
    out["charge"] = abs(c["charge"] - u.total_charge()) / qabs if qabs > 0 else abs(c["charge"])
    out["particles"] = int(c["particles"]) - u.num_particles()
    return out


def compare(u, sim, species, w=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """Every comparison of a GPU state with the oracle's: fields, per-cell particle sums,
    conservation sums."""
    return {"fields": field_errors(u, sim), "particles": particle_errors(u, sim, species, w, origin),
            "conservation": conservation_errors(u, sim, species)}
