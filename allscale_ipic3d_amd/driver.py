"""The reference's main driver (src/ipic3d.cxx:13-115) over the GPU step.

    python -m allscale_ipic3d_amd.driver <config.inp> [--out-dir DIR] [--cycles N]

reads a configuration file, sets the universe up as the driver does (ipic3d.cxx:41-85: time
step 0.01 s, the speed of light, cells scaled by the Earth radius, the planet at the coordinate
origin in the middle of the box, an external dipole of 3.07e-5 T), fills it through the
reference's own random streams (initCells over uniform_pos_normal_speed: positions uniform in
the middle 75% of the box, velocities normal with the speed of a 10 MeV proton as standard
deviation) and runs `ncycles` steps of the loop as shipped (simulator.h:181-231: the analytic
dipole mover with sub-cycling, reflection at the box faces, the planet as a sink; no deposit,
no field solve).  It writes <base>.out (outputNumberOfParticlesPerCell, cell.h:1360-1395) and
<base>ConservedQuantities.out (writeOutputData, simulator.h:73-91).

The field energies in that log are those of the initial field grid (initFields / getDipoleFieldAt,
field.h:43-110, summed by getFieldEnergy, field.h:443-459, with the accessors of
simulator.h:75-76), evaluated once on the host: the loop as shipped never changes the fields,
and the GPU's N-periodic grid does not hold the reference's (N+1)^3 ghost-bounded node set.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import FLAG_REFLECT, FLAG_REPRODUCIBLE, Simulation, desc_for, field_init, field_init_host, ref_distribution
from . import io as ipic_io

E_CHARGE = 1.602176565e-19   # elementary charge, C (ipic3d.cxx:70)
M_PROTON = 1.672621777e-27   # proton mass, kg (ipic3d.cxx:72)
EARTH_RADIUS = 6378137.0     # planet radius, m (ipic3d.cxx:54)
SPEED_OF_LIGHT = 299792458.0  # m/s (ipic3d.cxx:53)
DRIVER_DT = 0.01             # s (ipic3d.cxx:52)
EXTERNAL_BZ = 3.07e-5        # T (ipic3d.cxx:65)


class Parameters:
    """The keys of a configuration file the driver reads (parameters.h:145-380): `key = value`
    lines, `#` starts a comment, vector keys take one value per species."""

    SCALARS = {"dt": float, "ncycles": int, "Lx": float, "Ly": float, "Lz": float, "nxc": int, "nyc": int,
               "nzc": int, "B0x": float, "B0y": float, "B0z": float, "B1x": float, "B1y": float, "B1z": float,
               "FieldOutputCycle": int, "ParticlesOutputCycle": int}
    VECTORS = {"npcelx": int, "npcely": int, "npcelz": int, "u0": float, "v0": float, "w0": float}

    def __init__(self, path):
        vals = {}
        with open(path) as f:
            for line in f:
                body = line.split("#", 1)[0]
                if "=" not in body:
                    continue
                key, val = (t.strip() for t in body.split("=", 1))
                if key in self.SCALARS:
                    vals[key] = self.SCALARS[key](val.split()[-1])
                elif key in self.VECTORS:
                    vals[key] = [self.VECTORS[key](v) for v in val.split()]
        missing = [k for k in ("ncycles", "Lx", "Ly", "Lz", "nxc", "nyc", "nzc", "npcelx", "npcely", "npcelz")
                   if k not in vals]
        if missing:
            raise ValueError(f"{path}: missing keys {missing}")
        self.values = vals
        self.ncycles = vals["ncycles"]
        self.ncells = (vals["nxc"], vals["nyc"], vals["nzc"])
        self.L = (vals["Lx"], vals["Ly"], vals["Lz"])
        # parameters.h:376-378
        self.dspace = tuple(self.L[k] / self.ncells[k] for k in range(3))
        self.B0 = (vals.get("B0x", 0.0), vals.get("B0y", 0.0), vals.get("B0z", 0.0))
        self.B1 = (vals.get("B1x", 0.0), vals.get("B1y", 0.0), vals.get("B1z", 0.0))
        self.npcel = (vals["npcelx"], vals["npcely"], vals["npcelz"])
        self.drift = [(u, v, w) for u, v, w in zip(vals.get("u0", [0.0]), vals.get("v0", [0.0]), vals.get("w0", [0.0]))]
        self.field_output_cycle = vals.get("FieldOutputCycle", 100)
        self.particles_output_cycle = vals.get("ParticlesOutputCycle", 100)


@dataclass
class DriverCase:
    """The universe, initial distribution and run length ipic3d.cxx:41-85 derives from a
    configuration (all lengths in metres)."""
    base_name: str
    n: tuple
    cell_width: tuple
    origin: tuple
    object_center: tuple
    dt: float
    c: float
    planet_radius: float
    external_B: tuple
    B0: tuple                 # InitProperties::magneticField (init_properties.h:47)
    drift: tuple              # InitProperties::driftVelocity[0] (init_properties.h:40)
    num_particles: int
    pos_min: tuple
    pos_max: tuple
    v_mod: float
    ncycles: int
    field_output_cycle: int

    @property
    def dipole_moment(self) -> float:
        # magneticFieldTemp of moveParticles, cell.h:921
        return -self.external_B[2] * math.pow(self.planet_radius, 3)

    @property
    def species(self):
        # initCells sets q = e and qom = e / m_p for every particle (cell.h:644-645)
        return (E_CHARGE, E_CHARGE / M_PROTON)


def driver_case(params: Parameters, base_name: str = "") -> DriverCase:
    R = EARTH_RADIUS
    c = SPEED_OF_LIGHT
    n = tuple(int(v) for v in params.ncells)
    w = tuple(params.dspace[k] * R for k in range(3))            # cellWidth *= planetRadius
    size = tuple(w[k] * n[k] for k in range(3))                 # elementwiseProduct(cellWidth, size)
    center = (0.0, 0.0, 0.0)
    origin = tuple(center[k] - size[k] / 2.0 for k in range(3))
    ppc = params.npcel[0][0] * params.npcel[1][0] * params.npcel[2][0]
    num = n[0] * n[1] * n[2] * ppc
    K = 1e7 * E_CHARGE
    m = M_PROTON
    v_mod = c / math.sqrt(1.0 + (m * c * c) / K)
    low = tuple(origin[k] + 0.125 * size[k] for k in range(3))
    hig = tuple(low[k] + 0.75 * size[k] for k in range(3))
    return DriverCase(base_name=base_name, n=n, cell_width=w, origin=origin, object_center=center, dt=DRIVER_DT, c=c,
                      planet_radius=R, external_B=(0.0, 0.0, EXTERNAL_BZ), B0=tuple(params.B0),
                      drift=tuple(params.drift[0]), num_particles=num, pos_min=low, pos_max=hig, v_mod=v_mod,
                      ncycles=params.ncycles, field_output_cycle=params.field_output_cycle)


def case_field_init(case: DriverCase):
    """InitProperties / UniverseProperties of the driver's field (ipic3d.cxx:41-85): B0 and the
    drift of the configuration, the planet at the object centre, external B_z."""
    return field_init(drift=case.drift, B0=case.B0, object_center=case.object_center,
                      planet_radius=case.planet_radius, external_Bz=case.external_B[2])


def initial_field_energies(case: DriverCase):
    """getFieldEnergy of E and of B + Bext (simulator.h:75-79) over the nodes initFields fills
    (field.h:81-118), from the library's host initialisation (ipic_field_init_host)."""
    desc = desc_for(case.n, case.cell_width, case.origin)
    return field_init_host(desc, case_field_init(case))["energies"]


def make_simulation(case: DriverCase, tile=(0, 0, 0), flags=0) -> Simulation:
    """The universe of the case on the GPU with the loop as shipped: the analytic dipole mover,
    reflection at the box faces, the planet as a sink, the static field."""
    return Simulation(case.n, cell_width=case.cell_width, origin=case.origin, dt=case.dt, c=case.c,
                      species=(case.species,), field_substeps=0, sink_center=case.object_center,
                      sink_radius=case.planet_radius, tile=tile, flags=FLAG_REFLECT | flags,
                      dipole_moment=case.dipole_moment)


def case_distribution(case: DriverCase, nthreads=0):
    """uniform_pos_normal_speed<>(low, hig, 0, (v_mod, v_mod, v_mod)) of ipic3d.cxx:77-82."""
    return ref_distribution(case.pos_min, case.pos_max, (0.0, 0.0, 0.0), (case.v_mod,) * 3, case.num_particles,
                            seed=0, nthreads=nthreads)


def run(config, out_dir=".", ncycles=None, write=True, tile=(0, 0, 0), verbose=False, reproducible=False):
    """Runs a configuration as the reference's driver does; returns the final per-cell counts
    (row-major), the conserved-quantity rows and the timings.  reproducible: each cell's
    arrivals in the order of their contents (IPIC_FLAG_REPRODUCIBLE), so that two runs write
    identical files."""
    base = os.path.splitext(os.path.basename(config))[0]
    case = driver_case(Parameters(config), base)
    steps = case.ncycles if ncycles is None else int(ncycles)
    sim = make_simulation(case, tile=tile, flags=FLAG_REPRODUCIBLE if reproducible else 0)
    t0 = time.perf_counter()
    sim.init_reference(0, case_distribution(case))
    t_init = time.perf_counter() - t0
    e_energy, b_energy = initial_field_energies(case)
    rows = []
    log = None
    if write:
        os.makedirs(out_dir, exist_ok=True)
        if case.field_output_cycle > 0:
            log = ipic_io.ConservedLog(os.path.join(out_dir, base + "ConservedQuantities.out"), case.field_output_cycle)
    every = case.field_output_cycle
    ms = []
    for i in range(steps):
        if every > 0 and (i % every == 0 or i + 1 == steps):
            c = sim.conservation()
            rows.append((i, c["momentum"], e_energy, b_energy, c["kinetic_energy"]))
            if log:
                log.write(*rows[-1])
        st = sim.step(1)
        ms.append(1e3 * (st["first_step_s"] + st["remaining_steps_s"]))
    if log:
        log.close()
    counts = sim.cell_counts()
    if write:
        ipic_io.write_cell_counts(os.path.join(out_dir, base + ".out"), counts, case.n)
    removed = case.num_particles - int(counts.sum())
    if verbose:
        print(f"{case.num_particles} initial particles, {steps} cycles, init {t_init:.3f} s, "
              f"{removed} removed by the planet; first step {ms[0] if ms else 0:.3f} ms, "
              f"mean of the rest {np.mean(ms[1:]) if len(ms) > 1 else 0:.3f} ms", flush=True)
    sim.close()
    return {"case": case, "counts": counts, "rows": rows, "init_s": t_init, "step_ms": ms}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("config")
    ap.add_argument("--out-dir", default=".")
    ap.add_argument("--cycles", type=int, default=None)
    ap.add_argument("--reproducible", action="store_true",
                    help="order each cell's arrivals by their contents: runs are bitwise reproducible")
    a = ap.parse_args(argv)
    run(a.config, out_dir=a.out_dir, ncycles=a.cycles, verbose=True, reproducible=a.reproducible)
    return 0


if __name__ == "__main__":
    sys.exit(main())
