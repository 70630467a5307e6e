"""Shared helpers for the parity tests: identical inputs for the oracle and the GPU library,
and particle matching between the two (the GPU keeps a different within-cell order)."""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

import oracle as O

FOUR_PI = 16.0 * np.arctan(1.0)


def q_factor(qom, w, ppc, rho=1.0):
    """Macro-particle charge of cell.h:715-718: sign(qom) * vol / ppc * rho / (4 pi)."""
    return np.sign(qom) * (w[0] * w[1] * w[2]) / ppc * rho / FOUR_PI


def make_case(n=8, ppc=64, species=((-1.0, 0.02),), w=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
              dt=0.1, B=(0.0, 0.0, 1.0), E_sigma=0.0, seed=0, drift=(0.0, 0.0, 0.0), field_substeps=1,
              sink_center=None, sink_radius=0.0, ppc_per_cell=None):
    """Builds an oracle universe with the synthetic generator and returns it with the
    particles (per species) and fields to upload to the GPU."""
    if np.isscalar(n):
        n = (n, n, n)
    u = O.Universe(n, w=w, origin=origin, dt=dt, field_substeps=field_substeps,
                   sink_center=sink_center, sink_radius=sink_radius)
    ncell = n[0] * n[1] * n[2]
    counts = np.full(ncell, ppc, dtype=np.int64) if ppc_per_cell is None else np.asarray(ppc_per_cell, np.int64)
    sp = []
    parts = []
    for s, (qom, vth) in enumerate(species):
        q = q_factor(qom, w, max(ppc, 1))
        sp.append((q, qom))
        tmp = O.Universe(n, w=w, origin=origin, dt=dt)
        tmp.init_species(s, seed, counts, q, qom, (vth, vth, vth), drift)
        p, _ = tmp.particles()
        tmp.close()
        parts.append(p)
        u.add_particles(p)
    rng = np.random.default_rng(seed + 12345)
    E = rng.normal(0.0, E_sigma, size=(ncell, 3)) if E_sigma > 0 else np.zeros((ncell, 3))
    Bc = np.tile(np.asarray(B, dtype=np.float64), (ncell, 1))
    u.set_fields(E, Bc)
    return u, sp, parts, E, Bc


def gpu_sim_for(case, track_ids=True, **kw):
    import allscale_ipic3d_amd as G
    u, sp, parts, E, Bc = case
    p = u.props
    sim = G.Simulation(tuple(u.n), cell_width=tuple(p.w), origin=tuple(p.origin), dt=p.dt, c=p.c,
                       species=sp, field_substeps=p.field_substeps,
                       sink_center=tuple(p.sink_center) if p.sink else None, sink_radius=p.sink_radius,
                       track_ids=track_ids, **kw)
    for s, ps in enumerate(parts):
        sim.upload_particles(s, ps["pos"], ps["vel"])
    sim.upload_fields(E, Bc)
    return sim


def oracle_species(u, qom):
    parts, cell = u.particles()
    m = parts["qom"] == qom
    return parts[m], cell[m]


def match(ref_pos, ref_vel, got_pos, got_vel, box, vscale):
    """Nearest-neighbour matching of particles in (position/box-periodic, velocity) space.
    Returns the index into the reference for every GPU particle and the distances."""
    L = np.asarray(box, dtype=np.float64)
    big = 1e9
    a = np.concatenate([np.mod(ref_pos, L), ref_vel / vscale + big / 2], 1)
    b = np.concatenate([np.mod(got_pos, L), got_vel / vscale + big / 2], 1)
    boxsize = np.concatenate([L, [big * 4] * 3])
    tree = cKDTree(a, boxsize=boxsize)
    d, idx = tree.query(b, k=1)
    return idx, d


def periodic_delta(a, b, L):
    d = a - b
    return d - np.round(d / L) * L


def rel_err(got, ref):
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(np.asarray(got) - ref) / (den if den > 0 else 1.0))
