"""The uneven-load dipole configuration (a centred z dipole of moment 1 as the magnetic
background, drift 0.1 along x, density 8 (1 + 9 exp(-r^2 / 5^2)) particles per cell, a sink of
radius 2 at the centre: the dipole set up by initFields (field.h:43-138) and gathered, the
sub-cycled Boris mover of cell.h:896-964, the sink of cell.h:1123-1128) for the configuration's
200 steps at 32^3, against the FP64 oracle on the same inputs.  Per-cell particle sums, fields and the FP64
conservation sums are compared (oracle/compare.py) at steps 1, 50, 100 and 200, and every
checkpoint prints its errors, so a run shows how they grow.  The full-size run of the same
configuration (128^3, 74.9M particles, 200 steps) is test_full_size_invariants_config3."""
import json

import numpy as np
import pytest

import oracle.compare as OC
from tests._helpers import gpu_sim_for, make_case
from tests.test_gpu_parity import _dipole_init

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.timeout(900)
def test_config3_shape_200_steps_against_oracle():
    import allscale_ipic3d_amd as G
    n = (32, 32, 32)
    init, c = _dipole_init(n)
    fi = G.field_init_host(G.desc_for(n), init)
    x, y, z = np.meshgrid(*[np.arange(k) + 0.5 for k in n], indexing="ij")
    r2 = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2).reshape(-1)
    ppc = np.rint(8 * (1 + 9 * np.exp(-r2 / 25.0))).astype(np.int64)
    u, sp, parts, _, _ = make_case(n=n, ppc=8, species=((-1.0, 0.05),), dt=0.1, B=(0.0, 0.0, 0.0), seed=9,
                                   drift=(0.1, 0.0, 0.0), sink_center=tuple(c), sink_radius=2.0,
                                   ppc_per_cell=ppc)
    u.set_fields(fi["E"], fi["Bc"])
    sim = gpu_sim_for((u, sp, parts, fi["E"], fi["Bc"]), track_ids=False)
    assert sim.init_fields(init) == fi["energies"]
    npart = int(ppc.sum())
    done = 0
    for step in (1, 50, 100, 200):
        u.step(step - done)
        sim.step(step - done)
        done = step
        res = OC.compare(u, sim, sp)
        print(f"step {step}:", json.dumps(res), flush=True)
        for k, e in res["fields"].items():
            assert e <= TOL, f"step {step}: {k} rel error {e:.3e}"
        p = res["particles"][0]
        # a particle that FP32 and FP64 place on either side of a face or of the sink's sphere
        assert abs(p["particles"] - p["particles_ref"]) <= 1e-4 * npart
        assert p["flip_particles"] <= 1e-4 * npart
        assert p["offset_max"] <= TOL, f"step {step}: offset {p['offset_max']:.3e}"
        assert p["velocity_max"] <= TOL, f"step {step}: velocity {p['velocity_max']:.3e}"
        for k in ("kinetic_energy", "momentum", "field_energy_E", "field_energy_B"):
            assert res["conservation"][k] <= TOL, f"step {step}: {k} {res['conservation'][k]:.3e}"
    sim.close()
    u.close()
