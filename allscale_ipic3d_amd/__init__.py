"""MI355X-native particle-in-cell step of the AllScale iPIC3D pilot.

The compute path is the HIP library ``_lib/libipic_gpu.so`` (C ABI in include/ipic_gpu.h).
This module binds it with ctypes.  There is no CPU fallback: if the library is missing or
no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libipic_gpu.so")
MAX_SPECIES = 8

FLAG_TRACK_IDS = 1
FLAG_PHASE_TIMING = 4
FLAG_REFLECT = 8
FLAG_UNIFORM_BACKGROUND = 16
FLAG_DIPOLE_MOVER = 32
FLAG_FORWARD_SOLVER = 64
FLAG_REPRODUCIBLE = 128


class SpeciesDesc(C.Structure):
    _fields_ = [("q", C.c_double), ("qom", C.c_double)]


class Desc(C.Structure):
    _fields_ = [
        ("n", C.c_int32 * 3),
        ("nspecies", C.c_int32),
        ("cell_width", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("species", SpeciesDesc * MAX_SPECIES),
        ("field_substeps", C.c_int32),
        ("sink", C.c_int32),
        ("sink_center", C.c_double * 3),
        ("sink_radius", C.c_double),
        ("tile", C.c_int32 * 3),
        ("flags", C.c_uint32),
        ("dipole_moment", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("first_step_s", C.c_double),
        ("remaining_steps_s", C.c_double),
        ("steps", C.c_int64),
        ("particles", C.c_int64),
        ("removed", C.c_int64),
        ("ms_field", C.c_double),
        ("ms_push", C.c_double),
        ("ms_scan", C.c_double),
        ("ms_deposit", C.c_double),
        ("operator_steps", C.c_int64),
        ("fallback_steps", C.c_int64),
        ("audits", C.c_int64),
        ("max_audit_error", C.c_double),
        ("max_audit_error_increment", C.c_double),
        ("max_divb_residual", C.c_double),
        ("max_energy_residual", C.c_double),
        ("max_audit_error_evolving", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Conservation(C.Structure):
    _fields_ = [
        ("kinetic_energy", C.c_double),
        ("momentum", C.c_double),
        ("charge", C.c_double),
        ("field_energy_E", C.c_double),
        ("field_energy_B", C.c_double),
        ("divB_max", C.c_double),
        ("divB_l2", C.c_double),
        ("EdotJ", C.c_double),
        ("particles", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FieldOperatorDesc(C.Structure):
    _fields_ = [
        ("params", C.c_void_p),
        ("in_E", C.c_float), ("in_B", C.c_float), ("in_J", C.c_float),
        ("out_E", C.c_float), ("out_P", C.c_float),
        ("substeps", C.c_int32),
        ("audit_every", C.c_int32),
        ("divb_tol", C.c_double),
        ("energy_tol", C.c_double),
        ("audit_tol", C.c_double),
    ]


class FieldState(C.Structure):
    _fields_ = [
        ("E", C.c_void_p), ("Bc", C.c_void_p), ("J", C.c_void_p),
        ("E0", C.c_double * 3), ("B0", C.c_double * 3),
        ("uniform_background", C.c_int32),
    ]


class RefDistribution(C.Structure):
    """ipic_ref_distribution: distribution::uniform_pos_normal_speed (cell.h:405-440)."""
    _fields_ = [
        ("pos_min", C.c_double * 3), ("pos_max", C.c_double * 3),
        ("vel_mean", C.c_double * 3), ("vel_stddev", C.c_double * 3),
        ("num_particles", C.c_uint64), ("seed", C.c_uint32), ("nthreads", C.c_int32),
    ]


class RefLattice(C.Structure):
    """ipic_ref_lattice: the parameter-file initCells (cell.h:699-776)."""
    _fields_ = [
        ("ppc", C.c_int32 * 3), ("nthreads", C.c_int32),
        ("vel_thermal", C.c_double * 3), ("vel_drift", C.c_double * 3),
    ]


class FieldInit(C.Structure):
    """ipic_field_init: what initFields / getDipoleFieldAt (field.h:43-118) read."""
    _fields_ = [
        ("drift", C.c_double * 3), ("B0", C.c_double * 3), ("object_center", C.c_double * 3),
        ("planet_radius", C.c_double), ("external_Bz", C.c_double),
    ]


def field_init(drift=(0.0, 0.0, 0.0), B0=(0.0, 0.0, 0.0), object_center=(0.0, 0.0, 0.0), planet_radius=0.0,
               external_Bz=0.0):
    f = FieldInit()
    f.drift[:] = [float(v) for v in drift]
    f.B0[:] = [float(v) for v in B0]
    f.object_center[:] = [float(v) for v in object_center]
    f.planet_radius = float(planet_radius)
    f.external_Bz = float(external_Bz)
    return f


def dipole_field_init(moment, center, radius, B0=(0.0, 0.0, 0.0), drift=(0.0, 0.0, 0.0)):
    """The field of a z dipole of the given moment (B = moment (3xz, 3yz, 2z^2-x^2-y^2) / r^5
    outside `radius`, zero inside) in getDipoleFieldAt's parameters: -external_Bz radius^3 =
    moment."""
    return field_init(drift=drift, B0=B0, object_center=center, planet_radius=radius,
                      external_Bz=-float(moment) / float(radius) ** 3)


def ref_distribution(pos_min, pos_max, vel_mean, vel_stddev, num_particles, seed=0, nthreads=0):
    d = RefDistribution()
    d.pos_min[:] = [float(v) for v in pos_min]
    d.pos_max[:] = [float(v) for v in pos_max]
    d.vel_mean[:] = [float(v) for v in vel_mean]
    d.vel_stddev[:] = [float(v) for v in vel_stddev]
    d.num_particles = int(num_particles)
    d.seed = int(seed) & 0xFFFFFFFF
    d.nthreads = int(nthreads)
    return d


# every symbol declared in include/ipic_gpu.h
ABI_SYMBOLS = (
    "ipic_gpu_create", "ipic_gpu_destroy", "ipic_gpu_last_error", "ipic_gpu_upload_particles",
    "ipic_gpu_init_particles", "ipic_gpu_upload_fields", "ipic_gpu_step", "ipic_gpu_num_particles",
    "ipic_gpu_cell_counts", "ipic_gpu_download_particles", "ipic_gpu_download_fields",
    "ipic_gpu_conservation", "ipic_gpu_device_info", "ipic_gpu_field_step", "ipic_gpu_particle_step",
    "ipic_gpu_set_field_operator", "ipic_gpu_field_state", "ipic_ref_cell_counts", "ipic_ref_particles",
    "ipic_ref_last_error", "ipic_gpu_init_reference", "ipic_field_init_host", "ipic_field_init_last_error",
    "ipic_gpu_init_fields", "ipic_ref_lattice_particles", "ipic_ref_lattice_charge", "ipic_gpu_init_lattice",
)

_lib = None


class IpicError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Loads libipic_gpu.so and declares the C ABI; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise IpicError(f"HIP library not built: {path} (run python -m allscale_ipic3d_amd.build)")
    L = C.CDLL(path)
    P, I32, I64, U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    L.ipic_gpu_create.argtypes = [C.POINTER(Desc), C.POINTER(C.c_void_p)]
    L.ipic_gpu_destroy.argtypes = [P]
    L.ipic_gpu_last_error.restype = C.c_char_p
    L.ipic_gpu_last_error.argtypes = [P]
    L.ipic_gpu_upload_particles.argtypes = [P, I32, I64, P, P]
    L.ipic_gpu_init_particles.argtypes = [P, I32, P, I32, P, P, U64]
    L.ipic_gpu_upload_fields.argtypes = [P, P, P]
    L.ipic_gpu_step.argtypes = [P, I64, C.POINTER(StepStats)]
    L.ipic_gpu_num_particles.argtypes = [P, I32, C.POINTER(C.c_int64)]
    L.ipic_gpu_cell_counts.argtypes = [P, I32, P]
    L.ipic_gpu_download_particles.argtypes = [P, I32, P, P, P, P]
    L.ipic_gpu_download_fields.argtypes = [P, P, P, P, P]
    L.ipic_gpu_conservation.argtypes = [P, C.POINTER(Conservation)]
    L.ipic_gpu_device_info.argtypes = [C.c_char_p, I32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    L.ipic_gpu_field_step.argtypes = [P, I32, C.POINTER(StepStats)]
    L.ipic_gpu_particle_step.argtypes = [P, C.POINTER(StepStats)]
    L.ipic_gpu_set_field_operator.argtypes = [P, C.POINTER(FieldOperatorDesc)]
    L.ipic_gpu_field_state.argtypes = [P, C.POINTER(FieldState)]
    L.ipic_ref_cell_counts.argtypes = [C.POINTER(Desc), C.POINTER(RefDistribution), P]
    L.ipic_ref_particles.argtypes = [C.POINTER(Desc), C.POINTER(RefDistribution), P, P, P]
    L.ipic_ref_last_error.restype = C.c_char_p
    L.ipic_ref_last_error.argtypes = []
    L.ipic_gpu_init_reference.argtypes = [P, I32, C.POINTER(RefDistribution), P]
    L.ipic_field_init_host.argtypes = [C.POINTER(Desc), C.POINTER(FieldInit), P, P, P, P]
    L.ipic_field_init_last_error.restype = C.c_char_p
    L.ipic_field_init_last_error.argtypes = []
    L.ipic_gpu_init_fields.argtypes = [P, C.POINTER(FieldInit), P]
    L.ipic_ref_lattice_particles.argtypes = [C.POINTER(Desc), C.POINTER(RefLattice), P, P]
    L.ipic_gpu_init_lattice.argtypes = [P, I32, C.POINTER(RefLattice)]
    for name in ABI_SYMBOLS:
        if name not in ("ipic_gpu_last_error", "ipic_ref_last_error", "ipic_field_init_last_error",
                        "ipic_ref_lattice_charge"):
            getattr(L, name).restype = C.c_int
    L.ipic_ref_lattice_charge.restype = C.c_double
    L.ipic_ref_lattice_charge.argtypes = [P, P, C.c_double, C.c_double]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def reference_init(desc, dist, particles=True):
    """The reference's initCells streams on the host (ipic_ref_cell_counts / ipic_ref_particles):
    per-cell counts (row-major) and, with particles=True, positions and velocities (n x 3
    doubles, row-major cells, draw order within a cell).  No GPU is used."""
    L = load_library()
    ncell = int(desc.n[0]) * int(desc.n[1]) * int(desc.n[2])
    counts = np.zeros(ncell, dtype=np.int64)
    if L.ipic_ref_cell_counts(C.byref(desc), C.byref(dist), _ptr(counts)) != 0:
        raise IpicError(L.ipic_ref_last_error().decode())
    if not particles:
        return counts
    n = int(counts.sum())
    pos = np.zeros((n, 3), dtype=np.float64)
    vel = np.zeros((n, 3), dtype=np.float64)
    if L.ipic_ref_particles(C.byref(desc), C.byref(dist), _ptr(counts), _ptr(pos), _ptr(vel)) != 0:
        raise IpicError(L.ipic_ref_last_error().decode())
    return counts, pos, vel


def ref_lattice(ppc, vel_thermal, vel_drift=(0.0, 0.0, 0.0), nthreads=0):
    lat = RefLattice()
    lat.ppc[:] = [int(v) for v in ppc]
    lat.nthreads = int(nthreads)
    lat.vel_thermal[:] = [float(v) for v in vel_thermal]
    lat.vel_drift[:] = [float(v) for v in vel_drift]
    return lat


def lattice_particles(desc, lat):
    """The parameter-file initCells on the host (ipic_ref_lattice_particles): positions and
    velocities (n x 3 doubles, row-major cells, lattice order within a cell).  No GPU is used."""
    L = load_library()
    ncell = int(desc.n[0]) * int(desc.n[1]) * int(desc.n[2])
    n = ncell * int(lat.ppc[0]) * int(lat.ppc[1]) * int(lat.ppc[2])
    pos = np.zeros((n, 3), dtype=np.float64)
    vel = np.zeros((n, 3), dtype=np.float64)
    if L.ipic_ref_lattice_particles(C.byref(desc), C.byref(lat), _ptr(pos), _ptr(vel)) != 0:
        raise IpicError(L.ipic_ref_last_error().decode())
    return pos, vel


def lattice_charge(cell_width, ppc, qom, rho=1.0):
    """q of every particle of the parameter-file initCells (cell.h:711-716)."""
    L = load_library()
    return float(L.ipic_ref_lattice_charge(_ptr(_f64(cell_width)), _ptr(np.asarray(ppc, dtype=np.int32)),
                                           float(qom), float(rho)))


def field_init_host(desc, init):
    """initFields + initBcFields on the host (ipic_field_init_host): E and node B at the N^3
    nodes, Bc at the N^3 centres (row-major, n x 3 doubles) and the logged initial field
    energies (getFieldEnergy of E and of B + Bext over the (N+1)^3 nodes).  No GPU is used."""
    L = load_library()
    ncell = int(desc.n[0]) * int(desc.n[1]) * int(desc.n[2])
    E, Bn, Bc = (np.zeros((ncell, 3)) for _ in range(3))
    en = np.zeros(2)
    if L.ipic_field_init_host(C.byref(desc), C.byref(init), _ptr(E), _ptr(Bn), _ptr(Bc), _ptr(en)) != 0:
        raise IpicError(L.ipic_field_init_last_error().decode())
    return {"E": E, "Bn": Bn, "Bc": Bc, "energies": (float(en[0]), float(en[1]))}


def desc_for(n, cell_width=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """An ipic_desc of the grid geometry only (for the host-side set-up functions)."""
    if np.isscalar(n):
        n = (n, n, n)
    d = Desc()
    d.n[:] = [int(v) for v in n]
    d.nspecies = 1
    d.cell_width[:] = [float(v) for v in cell_width]
    d.origin[:] = [float(v) for v in origin]
    d.dt, d.c = 1.0, 1.0
    return d


def device_info():
    L = load_library()
    name = C.create_string_buffer(256)
    cu = C.c_int32()
    mem = C.c_int64()
    rc = L.ipic_gpu_device_info(name, 256, C.byref(cu), C.byref(mem))
    if rc != 0:
        raise IpicError(L.ipic_gpu_last_error(None).decode())
    return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": mem.value}


class Simulation:
    """A universe resident on the GPU (the Universe of universe.h:12-51 behind the C ABI)."""

    def __init__(self, n, cell_width=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), dt=0.1, c=1.0,
                 species=((-1.0, -1.0),), field_substeps=1, sink_center=None, sink_radius=0.0,
                 tile=(0, 0, 0), track_ids=False, flags=0, dipole_moment=None):
        L = load_library()
        if np.isscalar(n):
            n = (n, n, n)
        d = Desc()
        d.n[:] = [int(v) for v in n]
        d.nspecies = len(species)
        d.cell_width[:] = [float(v) for v in cell_width]
        d.origin[:] = [float(v) for v in origin]
        d.dt, d.c = float(dt), float(c)
        for i, (q, qom) in enumerate(species):
            d.species[i].q = float(q)
            d.species[i].qom = float(qom)
        d.field_substeps = int(field_substeps)
        if sink_center is not None:
            d.sink = 1
            d.sink_center[:] = [float(v) for v in sink_center]
            d.sink_radius = float(sink_radius)
        d.tile[:] = [int(v) for v in tile]
        d.flags = int(flags) | (FLAG_TRACK_IDS if track_ids else 0)
        if dipole_moment is not None:  # the mover as shipped: analytic dipole, E = 0
            d.flags |= FLAG_DIPOLE_MOVER
            d.dipole_moment = float(dipole_moment)
        self.desc = d
        self.n = tuple(int(v) for v in n)
        self.ncell = self.n[0] * self.n[1] * self.n[2]
        self.nspecies = len(species)
        self.track_ids = bool(track_ids)
        h = C.c_void_p()
        rc = L.ipic_gpu_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise IpicError(f"ipic_gpu_create failed ({rc}): {L.ipic_gpu_last_error(None).decode()}")
        self._h = h
        self._L = L

    def _check(self, rc):
        if rc != 0:
            raise IpicError(f"libipic_gpu error {rc}: {self._L.ipic_gpu_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.ipic_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state in
    def upload_particles(self, species, pos, vel):
        pos = _f64(pos, (-1, 3))
        vel = _f64(vel, (-1, 3))
        assert len(pos) == len(vel)
        self._check(self._L.ipic_gpu_upload_particles(self._h, int(species), len(pos), _ptr(pos), _ptr(vel)))

    def init_particles(self, species, ppc, vth, drift=(0.0, 0.0, 0.0), seed=0):
        if np.isscalar(ppc):
            arr, uni = None, int(ppc)
        else:
            arr = np.ascontiguousarray(np.asarray(ppc, dtype=np.int32).reshape(-1))
            assert arr.size == self.ncell
            uni = 0
        self._check(self._L.ipic_gpu_init_particles(self._h, int(species), _ptr(arr), uni, _ptr(_f64(vth)),
                                                    _ptr(_f64(drift)), int(seed)))

    def init_reference(self, species, dist):
        """initCells with the reference's random streams (ipic_gpu_init_reference); returns the
        phase-1 per-cell counts."""
        counts = np.zeros(self.ncell, dtype=np.int64)
        self._check(self._L.ipic_gpu_init_reference(self._h, int(species), C.byref(dist), _ptr(counts)))
        return counts

    def init_lattice(self, species, lat):
        """The parameter-file initCells (cell.h:699-776) into `species` (ipic_gpu_init_lattice); the
        species' q should be lattice_charge(...) of the same lattice."""
        self._check(self._L.ipic_gpu_init_lattice(self._h, int(species), C.byref(lat)))

    def init_fields(self, init):
        """initFields + initBcFields (field.h:81-138) as the field state (ipic_gpu_init_fields);
        returns the logged initial field energies (E, B + Bext)."""
        en = np.zeros(2)
        self._check(self._L.ipic_gpu_init_fields(self._h, C.byref(init), _ptr(en)))
        return float(en[0]), float(en[1])

    def upload_fields(self, E=None, Bc=None):
        e = None if E is None else _f64(E, (self.ncell, 3))
        b = None if Bc is None else _f64(Bc, (self.ncell, 3))
        self._check(self._L.ipic_gpu_upload_fields(self._h, _ptr(e), _ptr(b)))

    # ---- stepping
    def step(self, nsteps=1):
        st = StepStats()
        self._check(self._L.ipic_gpu_step(self._h, int(nsteps), C.byref(st)))
        return st.as_dict()

    def field_step(self, use_operator=False, stats=None):
        """One field phase (leapfrog substeps, or the field operator); returns the stats."""
        st = stats if stats is not None else StepStats()
        self._check(self._L.ipic_gpu_field_step(self._h, 1 if use_operator else 0, C.byref(st)))
        return st

    def particle_step(self, stats=None):
        st = stats if stats is not None else StepStats()
        self._check(self._L.ipic_gpu_particle_step(self._h, C.byref(st)))
        return st

    def set_field_operator(self, params=None, scales=None, substeps=None, audit_every=0, divb_tol=1e-4,
                           energy_tol=1e-3, audit_tol=1e-3):
        """Installs the learned field step (params: FP32 vector, scales: dict with in_E, in_B,
        in_J, out_E, out_P); params=None switches it off."""
        if params is None:
            self._check(self._L.ipic_gpu_set_field_operator(self._h, None))
            return
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32).reshape(-1))
        d = FieldOperatorDesc()
        d.params = p.ctypes.data_as(C.c_void_p)
        for k in ("in_E", "in_B", "in_J", "out_E", "out_P"):
            setattr(d, k, float(scales[k]))
        d.substeps = int(self.desc.field_substeps if substeps is None else substeps)
        d.audit_every = int(audit_every)
        d.divb_tol, d.energy_tol, d.audit_tol = float(divb_tol), float(energy_tol), float(audit_tol)
        self._check(self._L.ipic_gpu_set_field_operator(self._h, C.byref(d)))

    def field_state(self):
        """Device pointers of dE, dBc, J (float4 x n^3) and the uniform background."""
        fs = FieldState()
        self._check(self._L.ipic_gpu_field_state(self._h, C.byref(fs)))
        return {"E": fs.E, "Bc": fs.Bc, "J": fs.J, "E0": tuple(fs.E0), "B0": tuple(fs.B0),
                "uniform_background": bool(fs.uniform_background)}

    # ---- state out
    def num_particles(self, species=-1):
        n = C.c_int64()
        self._check(self._L.ipic_gpu_num_particles(self._h, int(species), C.byref(n)))
        return n.value

    def cell_counts(self, species=-1):
        out = np.zeros(self.ncell, dtype=np.int64)
        self._check(self._L.ipic_gpu_cell_counts(self._h, int(species), _ptr(out)))
        return out

    def download_particles(self, species):
        n = self.num_particles(species)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        cell = np.zeros(n, dtype=np.int64)
        ids = np.zeros(n, dtype=np.uint32) if self.track_ids else None
        self._check(self._L.ipic_gpu_download_particles(self._h, int(species), _ptr(pos), _ptr(vel), _ptr(cell),
                                                        _ptr(ids)))
        out = {"pos": pos, "vel": vel, "cell": cell}
        if ids is not None:
            out["ids"] = ids
        return out

    def download_fields(self):
        E, Bc, Bn, J = (np.zeros((self.ncell, 3)) for _ in range(4))
        self._check(self._L.ipic_gpu_download_fields(self._h, _ptr(E), _ptr(Bc), _ptr(Bn), _ptr(J)))
        return {"E": E, "Bc": Bc, "Bn": Bn, "J": J}

    def conservation(self):
        c = Conservation()
        self._check(self._L.ipic_gpu_conservation(self._h, C.byref(c)))
        return c.as_dict()
