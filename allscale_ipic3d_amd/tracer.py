"""The particle tracer of src/pic_tracer.cxx on the GPU.

traceParticle (pic_tracer.cxx:77-148) follows every particle through the static field of a
z dipole at the coordinate origin (E = 0, sub-cycled Boris, as the shipped mover of
cell.h:919-945), wraps positions periodically, stops particles that enter the planet of
radius R at objectCenter, and every frame_interval steps counts particles per cell.  Here
all particles advance together in the GPU step with the dipole mover, the static field
(field_substeps 0), periodic wrap and the planet sink; the frames are the per-cell counts
at the start of steps 0, frame_interval, 2 frame_interval, ...  (The reference registers a
particle of such a step only if it survives the step; a particle that enters the planet in
that very step is counted here.)
"""
from __future__ import annotations

import numpy as np

from . import Simulation


def trace(positions, velocities, q, qom, n, cell_width, origin, dt, planet_radius, object_center, B1z, steps,
          frame_interval, c=1.0):
    """Returns (frames [nframes][ncell] int64, particles started).  positions are relative to
    the object centre, as the tracer's generator produces them (pic_tracer.cxx:86)."""
    n = tuple(int(v) for v in n)
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3) + np.asarray(object_center, dtype=np.float64)
    vel = np.asarray(velocities, dtype=np.float64).reshape(-1, 3)
    low = np.asarray(origin, dtype=np.float64)
    high = low + np.asarray(cell_width, dtype=np.float64) * np.asarray(n)
    # pic_tracer.cxx:88-97: no particle inside the planet or outside the universe
    d = pos - np.asarray(object_center, dtype=np.float64)
    keep = np.sum(d * d, axis=1) >= planet_radius ** 2
    keep &= np.all((pos <= high) & (pos >= low), axis=1)
    pos, vel = pos[keep], vel[keep]
    sim = Simulation(n, cell_width=cell_width, origin=origin, dt=dt, c=c, species=((q, qom),), field_substeps=0,
                     sink_center=tuple(object_center), sink_radius=planet_radius,
                     dipole_moment=-B1z * planet_radius ** 3)
    try:
        sim.upload_particles(0, pos, vel)
        frames = []
        for t in range(steps):
            if t % frame_interval == 0:
                frames.append(sim.cell_counts())
            sim.step(1)
        return np.asarray(frames), int(len(pos))
    finally:
        sim.close()
