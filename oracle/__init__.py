"""ctypes wrapper of the FP64 CPU oracle (oracle/ipic_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use this module as the checker.  Nothing in the product package imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libipic_oracle.so")

PARTICLE_DTYPE = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("q", "<f8"), ("qom", "<f8")])


class Props(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("w", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("field_substeps", C.c_int32),
        ("sink", C.c_int32),
        ("sink_center", C.c_double * 3),
        ("sink_radius", C.c_double),
        ("bc", C.c_int32),
        ("mover", C.c_int32),
        ("dipole_moment", C.c_double),
        ("solver", C.c_int32),
        ("pad_", C.c_int32),
    ]


class RefDist(C.Structure):
    _fields_ = [
        ("pos_min", C.c_double * 3), ("pos_max", C.c_double * 3),
        ("vel_mean", C.c_double * 3), ("vel_stddev", C.c_double * 3),
        ("num_particles", C.c_uint64), ("seed", C.c_uint32), ("pad_", C.c_int32),
        ("q", C.c_double), ("qom", C.c_double),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, D, I64, I32, U64 = C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int32, C.c_uint64
        L.or_create.restype = P
        L.or_create.argtypes = [C.POINTER(Props)]
        L.or_destroy.argtypes = [P]
        L.or_num_cells.restype = I64
        L.or_num_cells.argtypes = [P]
        L.or_add_particles.argtypes = [P, I64, P]
        L.or_num_particles.restype = I64
        L.or_num_particles.argtypes = [P]
        L.or_get_particles.argtypes = [P, P, P]
        L.or_cell_counts.argtypes = [P, P]
        L.or_set_fields.argtypes = [P, P, P]
        L.or_get_fields.argtypes = [P, P, P, P, P]
        L.or_set_J.argtypes = [P, P]
        for f in ("or_deposit", "or_field_solve", "or_move"):
            getattr(L, f).argtypes = [P, C.c_int]
        L.or_rebin.restype = I64
        L.or_rebin.argtypes = [P, C.c_int]
        L.or_step.argtypes = [P, I64, C.c_int]
        for f in ("or_kinetic_energy", "or_momentum", "or_total_charge"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [P]
        L.or_field_energy.argtypes = [P, D, D]
        L.or_divB.restype = C.c_double
        L.or_divB.argtypes = [P, D]
        L.or_ref_field_energy.restype = C.c_double
        L.or_ref_field_energy.argtypes = [P, C.c_int]
        L.or_update_position.argtypes = [P, C.c_double]
        L.or_update_velocity.argtypes = [P, P, P, C.c_double]
        L.or_trilinear.restype = C.c_double
        L.or_trilinear.argtypes = [P, P, C.c_double]
        L.or_ref_curlB.argtypes = [P, P, P, P, P]
        L.or_ref_curlE.argtypes = [P, P, P, P, P]
        L.or_ref_interpC2N.argtypes = [P, P, P, P, P]
        L.or_ref_interpN2C.argtypes = [P, P, P, P, P]
        L.or_ref_update_boundaries.argtypes = [I64, P, P, I64, P]
        L.or_ref_leapfrog_node.argtypes = [P, P, P, P, P, P, P, P, C.c_double, C.c_double]
        L.or_dipole_field.argtypes = [P, P, P, P, C.c_double, C.c_double, P, P, P]
        L.or_mix64.restype = U64
        L.or_mix64.argtypes = [U64]
        L.or_draw.argtypes = [U64, I32, I64, I64, P, P]
        L.or_init_species.argtypes = [P, I32, U64, P, C.c_double, C.c_double, P, P]
        L.or_ref_init.argtypes = [P, C.POINTER(RefDist), P, P]
        L.or_ref_normal_draws.argtypes = [U64, I64, P]
        L.or_ref_init_lattice.argtypes = [P, P, C.c_double, C.c_double, P, P, P]
        L.or_ref_uniform_draws.argtypes = [C.c_uint32, C.c_double, C.c_double, I64, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def ref_init(universe, pos_min, pos_max, vel_mean, vel_stddev, num_particles, q, qom, seed=0,
             particles=True):
    """initCells (cell.h:512-654) of distribution::uniform_pos_normal_speed with the reference's
    random streams (oracle/ref_init.cpp): per-cell counts and, with particles=True, the
    particles (PARTICLE_DTYPE, row-major cells) with q and qom set as initCells sets them."""
    d = RefDist()
    d.pos_min[:] = [float(v) for v in pos_min]
    d.pos_max[:] = [float(v) for v in pos_max]
    d.vel_mean[:] = [float(v) for v in vel_mean]
    d.vel_stddev[:] = [float(v) for v in vel_stddev]
    d.num_particles = int(num_particles)
    d.seed = int(seed) & 0xFFFFFFFF
    d.q, d.qom = float(q), float(qom)
    counts = np.zeros(universe.ncell, dtype=np.int64)
    rc = lib().or_ref_init(C.byref(universe.props), C.byref(d), _p(counts), None)
    if rc != 0:
        raise ValueError("initCells asserts for this particle count (no remainder to spread)")
    if not particles:
        return counts
    out = np.zeros(int(counts.sum()), dtype=PARTICLE_DTYPE)
    lib().or_ref_init(C.byref(universe.props), C.byref(d), _p(counts), _p(out))
    return counts, out


def ref_init_lattice(universe, ppc, qom, uth, u0, rho=1.0):
    """The parameter-file initCells (cell.h:699-776): lattice positions and Box-Muller
    Maxwellian velocities from the reference's per-cell std::minstd_rand streams
    (PARTICLE_DTYPE, row-major cells)."""
    p = i64(ppc)
    out = np.zeros(universe.ncell * int(np.prod(p)), dtype=PARTICLE_DTYPE)
    lib().or_ref_init_lattice(C.byref(universe.props), _p(p), float(qom), float(rho), _p(f64(uth)), _p(f64(u0)),
                              _p(out))
    return out


def ref_normal_draws(seed, n):
    out = np.zeros(int(n), dtype=np.float32)
    lib().or_ref_normal_draws(int(seed), int(n), _p(out))
    return out


def ref_uniform_draws(seed, a, b, n):
    out = np.zeros(int(n), dtype=np.float64)
    lib().or_ref_uniform_draws(int(seed) & 0xFFFFFFFF, float(a), float(b), int(n), _p(out))
    return out


class Universe:
    """N-periodic FP64 universe of the oracle (cells, E/Bc/Bn/J grids)."""

    def __init__(self, n, w=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), dt=0.1, c=1.0,
                 field_substeps=1, sink_center=None, sink_radius=0.0, bc="periodic", mover="gather",
                 dipole_moment=0.0, solver="leapfrog"):
        if np.isscalar(n):
            n = (n, n, n)
        p = Props()
        p.n[:] = [int(v) for v in n]
        p.w[:] = [float(v) for v in w]
        p.origin[:] = [float(v) for v in origin]
        p.dt, p.c, p.field_substeps = float(dt), float(c), int(field_substeps)
        if sink_center is not None:
            p.sink = 1
            p.sink_center[:] = [float(v) for v in sink_center]
            p.sink_radius = float(sink_radius)
        p.bc = {"periodic": 0, "reflect": 1}[bc]
        p.mover = {"gather": 0, "dipole": 1}[mover]
        p.dipole_moment = float(dipole_moment)
        p.solver = {"leapfrog": 0, "forward": 1}[solver]
        self.props = p
        self.n = tuple(int(v) for v in n)
        self.ncell = self.n[0] * self.n[1] * self.n[2]
        self._h = lib().or_create(C.byref(p))
        if not self._h:
            raise ValueError("invalid universe properties")

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # particles
    def add_particles(self, parts: np.ndarray):
        parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
        lib().or_add_particles(self._h, len(parts), _p(parts))

    def init_species(self, species, seed, counts, q, qom, vth, drift=(0.0, 0.0, 0.0)):
        counts = i64(np.broadcast_to(np.asarray(counts, dtype=np.int64), (self.ncell,)).copy())
        lib().or_init_species(self._h, int(species), int(seed), _p(counts), float(q), float(qom),
                              _p(f64(vth)), _p(f64(drift)))

    def num_particles(self) -> int:
        return int(lib().or_num_particles(self._h))

    def particles(self):
        n = self.num_particles()
        out = np.zeros(n, dtype=PARTICLE_DTYPE)
        cell = np.zeros(n, dtype=np.int64)
        lib().or_get_particles(self._h, _p(out), _p(cell))
        return out, cell

    def cell_counts(self) -> np.ndarray:
        c = np.zeros(self.ncell, dtype=np.int64)
        lib().or_cell_counts(self._h, _p(c))
        return c

    # fields ([ncell,3], index (x*ny+y)*nz+z)
    def set_fields(self, E=None, Bc=None):
        e = f64(E).reshape(-1, 3) if E is not None else None
        b = f64(Bc).reshape(-1, 3) if Bc is not None else None
        lib().or_set_fields(self._h, _p(e) if e is not None else None, _p(b) if b is not None else None)

    def set_J(self, J):
        j = f64(J).reshape(-1, 3)
        lib().or_set_J(self._h, _p(j))

    def fields(self):
        E, Bc, Bn, J = (np.zeros((self.ncell, 3)) for _ in range(4))
        lib().or_get_fields(self._h, _p(E), _p(Bc), _p(Bn), _p(J))
        return {"E": E, "Bc": Bc, "Bn": Bn, "J": J}

    # phases
    def deposit(self, nthreads=0):
        lib().or_deposit(self._h, nthreads)

    def field_solve(self, nthreads=0):
        lib().or_field_solve(self._h, nthreads)

    def move(self, nthreads=0):
        lib().or_move(self._h, nthreads)

    def rebin(self, nthreads=0) -> int:
        return int(lib().or_rebin(self._h, nthreads))

    def step(self, nsteps=1, nthreads=0):
        lib().or_step(self._h, int(nsteps), nthreads)

    # reductions
    def kinetic_energy(self):
        return lib().or_kinetic_energy(self._h)

    def momentum(self):
        return lib().or_momentum(self._h)

    def total_charge(self):
        return lib().or_total_charge(self._h)

    def field_energy(self):
        a, b = C.c_double(), C.c_double()
        lib().or_field_energy(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def divB(self):
        l2 = C.c_double()
        mx = lib().or_divB(self._h, C.byref(l2))
        return mx, l2.value

    def ref_field_energy(self, which):
        return lib().or_ref_field_energy(self._h, int(which))


def make_particles(pos, vel, q, qom) -> np.ndarray:
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    vel = np.atleast_2d(np.asarray(vel, dtype=np.float64))
    out = np.zeros(len(pos), dtype=PARTICLE_DTYPE)
    out["pos"] = pos
    out["vel"] = vel
    out["q"] = q
    out["qom"] = qom
    return out


def boris(vel, E, B, qom, dt):
    p = make_particles([[0.0, 0.0, 0.0]], [vel], 1.0, qom)
    lib().or_update_velocity(_p(p), _p(f64(E)), _p(f64(B)), float(dt))
    return p["vel"][0].copy()


def draw(seed, species, cell, index):
    u = np.zeros(3)
    n = np.zeros(3)
    lib().or_draw(int(seed), int(species), int(cell), int(index), _p(u), _p(n))
    return u, n
