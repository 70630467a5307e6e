"""Learned field step: training data from the GPU simulator, training and evaluation of the
3D-convolutional operator of include/ipic_nop.h, all through the HIP kernels of
libipic_gpu.so.  Device memory is handled through the library's own plumbing calls
(ipic_dev_*), so no GPU array library is needed.

The operator stands in for `substeps` leapfrog substeps of the field phase (field.h:356-372
with J held fixed).  Training pairs are (dE, dBc, J) before a field phase and the increments
(E' - E, Bc' - Bc) that plain leapfrog produces over it, taken from simulator runs.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import IpicError, Simulation, load_library

NOP_SYMBOLS = (
    "ipic_nop_last_error", "ipic_nop_param_count", "ipic_nop_layer", "ipic_nop_workspace_bytes",
    "ipic_nop_prepare", "ipic_nop_pack", "ipic_nop_forward", "ipic_nop_apply", "ipic_nop_loss",
    "ipic_nop_backward", "ipic_nop_adam", "ipic_dev_alloc", "ipic_dev_free", "ipic_dev_copy",
    "ipic_dev_zero", "ipic_dev_sync", "ipic_dev_sub4", "ipic_nop_apply_noalloc",
)

LAYERS = [(16, 32)] + [(32, 32)] * 8 + [(32, 16)]
REAL_IO = [(9, 32)] + [(32, 32)] * 8 + [(32, 6)]
FLOP_PER_CELL = 2 * 27 * (9 * 32 + 8 * 32 * 32 + 32 * 6)   # real channels, one inference


class Scales(C.Structure):
    _fields_ = [("in_E", C.c_float), ("in_B", C.c_float), ("in_J", C.c_float),
                ("out_E", C.c_float), ("out_P", C.c_float)]


_nop = None


def nop_lib():
    """libipic_gpu.so with the operator entry points declared."""
    global _nop
    if _nop is not None:
        return _nop
    L = load_library()
    P, I32, I64, F32, F64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
    N3 = C.POINTER(C.c_int32)
    S = C.POINTER(Scales)
    F3 = C.POINTER(C.c_float)
    L.ipic_nop_last_error.restype = C.c_char_p
    L.ipic_nop_last_error.argtypes = []
    L.ipic_nop_param_count.restype = I64
    L.ipic_nop_param_count.argtypes = []
    L.ipic_nop_layer.argtypes = [I32, C.POINTER(I64), C.POINTER(I64), C.POINTER(I32), C.POINTER(I32)]
    L.ipic_nop_workspace_bytes.restype = I64
    L.ipic_nop_workspace_bytes.argtypes = [N3, I32]
    L.ipic_nop_prepare.argtypes = [P, P, P, P]
    L.ipic_nop_pack.argtypes = [N3, P, P, P, S, P, P]
    L.ipic_nop_forward.argtypes = [N3, P, P, P, P, I32, P, P]
    L.ipic_nop_apply.argtypes = [N3, P, S, F3, P, P, P, P]
    L.ipic_nop_apply_noalloc.argtypes = [N3, P, S, F3, P, P, P]
    L.ipic_nop_loss.argtypes = [N3, P, S, F3, P, P, F32, F32, P, P, P, P, P]
    L.ipic_nop_backward.argtypes = [N3, P, P, P, P, P, P, P]
    L.ipic_nop_adam.argtypes = [P, P, P, P, I64, F32, F32, F32, F32, P, P, P]
    L.ipic_dev_alloc.argtypes = [I64, C.POINTER(C.c_void_p)]
    L.ipic_dev_free.argtypes = [P]
    L.ipic_dev_copy.argtypes = [P, P, I64]
    L.ipic_dev_zero.argtypes = [P, I64, P]
    L.ipic_dev_sync.argtypes = [P]
    L.ipic_dev_sub4.argtypes = [P, P, P, I64, P]
    for name in NOP_SYMBOLS:
        if name not in ("ipic_nop_last_error", "ipic_nop_param_count", "ipic_nop_workspace_bytes"):
            getattr(L, name).restype = C.c_int
    _nop = L
    return L


def _chk(rc):
    if rc != 0:
        raise IpicError(f"libipic_gpu operator error {rc}: {nop_lib().ipic_nop_last_error().decode()}")


def param_count():
    return int(nop_lib().ipic_nop_param_count())


def _n3(n):
    return (C.c_int32 * 3)(*[int(v) for v in n])


class Dev:
    """A device allocation."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _chk(nop_lib().ipic_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    def free(self):
        if getattr(self, "ptr", None):
            nop_lib().ipic_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _chk(nop_lib().ipic_dev_copy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(count, dtype=dtype)
        _chk(nop_lib().ipic_dev_copy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def zero(self):
        _chk(nop_lib().ipic_dev_zero(self.ptr, self.nbytes, None))


def dev_copy(dst_ptr, src_ptr, nbytes):
    _chk(nop_lib().ipic_dev_copy(dst_ptr, src_ptr, int(nbytes)))


def init_params(seed=0, out_gain=0.1, res_gain=0.1):
    """He-normal weights over the real channels (padding zero), zero biases; the last conv of
    each residual block and the projection start small."""
    L = nop_lib()
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), dtype=np.float32)
    for l, (rci, rco) in enumerate(REAL_IO):
        wo, bo, ci, co = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        _chk(L.ipic_nop_layer(l, C.byref(wo), C.byref(bo), C.byref(ci), C.byref(co)))
        std = np.sqrt(2.0 / (27 * rci))
        if l == len(REAL_IO) - 1:
            std *= out_gain
        elif l >= 1 and l % 2 == 0:
            std *= res_gain
        w = np.zeros((27, co.value, ci.value), dtype=np.float32)
        w[:, :rco, :rci] = rng.normal(0.0, std, size=(27, rco, rci))
        p[wo.value:bo.value] = w.reshape(-1)
    return p


class Network:
    """Parameters, BF16 weight copies and work buffers of the operator on an n grid."""

    def __init__(self, n, scales, params=None, train=True, stream=None):
        L = nop_lib()
        self.n = tuple(int(v) for v in n)
        self.V = self.n[0] * self.n[1] * self.n[2]
        self.n3 = _n3(self.n)
        self.scales = dict(scales)
        self.cs = Scales(*[float(self.scales[k]) for k in ("in_E", "in_B", "in_J", "out_E", "out_P")])
        self.stream = stream
        self.train = bool(train)
        self.np_ = param_count()
        p = init_params() if params is None else np.asarray(params, dtype=np.float32)
        assert p.size == self.np_
        self.params = Dev(4 * self.np_)
        self.params.upload(p)
        self.wf = Dev(2 * self.np_)
        self.wb = Dev(2 * self.np_) if train else None
        wsb = L.ipic_nop_workspace_bytes(self.n3, 1 if train else 0)
        if wsb < 0:
            _chk(-1)
        self.ws = Dev(wsb)
        self.out = Dev(4 * 16 * self.V)
        self.u = Dev(2 * 16 * self.V)
        if train:
            self.grad = Dev(4 * self.np_)
            self.m = Dev(4 * self.np_)
            self.v = Dev(4 * self.np_)
            self.m.zero()
            self.v.zero()
            self.dout = Dev(2 * 16 * self.V)
            self.g4 = Dev(16 * self.V)
            self.psi4 = Dev(16 * self.V)
            self.loss = Dev(8)
            self.step_count = 0
        self.prepare()

    def prepare(self):
        _chk(nop_lib().ipic_nop_prepare(self.params.ptr, self.wf.ptr, self.wb.ptr if self.wb else None, self.stream))

    def get_params(self):
        return self.params.download(np.float32, self.np_)

    def set_params(self, p):
        self.params.upload(np.asarray(p, dtype=np.float32))
        self.prepare()

    def pack(self, dE_ptr, dBc_ptr, J_ptr, u_ptr=None):
        _chk(nop_lib().ipic_nop_pack(self.n3, dE_ptr, dBc_ptr, J_ptr, C.byref(self.cs),
                                     u_ptr if u_ptr is not None else self.u.ptr, self.stream))

    def forward(self, u_ptr=None, train=False):
        _chk(nop_lib().ipic_nop_forward(self.n3, u_ptr if u_ptr is not None else self.u.ptr, self.wf.ptr,
                                        self.params.ptr, self.ws.ptr, 1 if train else 0, self.out.ptr, self.stream))

    def iw(self, cell_width=(1.0, 1.0, 1.0)):
        return (C.c_float * 3)(*[1.0 / float(w) for w in cell_width])

    def loss_and_grad(self, tE_ptr, tB_ptr, varE, varB, u_ptr=None, cell_width=(1.0, 1.0, 1.0)):
        """Forward (train), loss and backward; the gradient is left in self.grad.  Returns L."""
        L = nop_lib()
        self.grad.zero()
        self.loss.zero()
        self.forward(u_ptr, train=True)
        _chk(L.ipic_nop_loss(self.n3, self.out.ptr, C.byref(self.cs), self.iw(cell_width), tE_ptr, tB_ptr,
                             float(varE), float(varB), self.loss.ptr, self.dout.ptr, self.g4.ptr, self.psi4.ptr,
                             self.stream))
        _chk(L.ipic_nop_backward(self.n3, u_ptr if u_ptr is not None else self.u.ptr, self.wb.ptr, self.params.ptr,
                                 self.ws.ptr, self.dout.ptr, self.grad.ptr, self.stream))
        return float(self.loss.download(np.float64, 1)[0])

    def eval_loss(self, tE_ptr, tB_ptr, varE, varB, u_ptr=None, cell_width=(1.0, 1.0, 1.0)):
        L = nop_lib()
        self.loss.zero()
        self.forward(u_ptr, train=False)
        _chk(L.ipic_nop_loss(self.n3, self.out.ptr, C.byref(self.cs), self.iw(cell_width), tE_ptr, tB_ptr,
                             float(varE), float(varB), self.loss.ptr, self.dout.ptr, self.g4.ptr, self.psi4.ptr,
                             self.stream))
        return float(self.loss.download(np.float64, 1)[0])

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        _chk(nop_lib().ipic_nop_adam(self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.step_count,
                                     float(lr), float(beta1), float(beta2), float(eps), self.wf.ptr, self.wb.ptr,
                                     self.stream))

    def apply(self, dE_ptr, dBc_ptr, psi_ptr, cell_width=(1.0, 1.0, 1.0)):
        _chk(nop_lib().ipic_nop_apply(self.n3, self.out.ptr, C.byref(self.cs), self.iw(cell_width), dE_ptr, dBc_ptr,
                                      psi_ptr, self.stream))

    def sync(self):
        _chk(nop_lib().ipic_dev_sync(self.stream))


class Snapshots:
    """Training pairs on the device: (dE, dBc, J) before a field phase and the leapfrog
    increments (tE, tB) over it, float4 x V each."""

    def __init__(self, n, count):
        self.n = tuple(int(v) for v in n)
        self.V = self.n[0] * self.n[1] * self.n[2]
        self.count = int(count)
        self.size = 16 * self.V
        self.buf = Dev(5 * self.size * self.count)
        self.filled = 0
        self.meta = []

    def ptrs(self, i):
        base = self.buf.ptr + 5 * self.size * i
        return {k: base + j * self.size for j, k in enumerate(("E", "Bc", "J", "tE", "tB"))}

    def download(self, i):
        """Host copy of pair i as [V][3] float32 arrays."""
        out = {}
        for k, p in self.ptrs(i).items():
            a = np.empty((self.V, 4), dtype=np.float32)
            dev_copy(a.ctypes.data, p, a.nbytes)
            out[k] = a[:, :3].copy()
        return out


def config2_simulation(n, ppc, seed, field_substeps=8, dt=0.05, sigma_E=1e-3, B0=(0.0, 0.2, 1.0), flags=0):
    """The two-species configuration: electrons (q/m -1, vth 0.05) and ions (q/m 1/64,
    vth 0.006), ppc/2 each per cell, B = B0, E ~ N(0, sigma_E^2) per node component."""
    half = ppc // 2
    q = 1.0 / (half * 4.0 * np.pi)
    from . import FLAG_UNIFORM_BACKGROUND
    sim = Simulation(n, dt=dt, c=1.0, species=((-q, -1.0), (q, 1.0 / 64)), field_substeps=field_substeps,
                     flags=flags | FLAG_UNIFORM_BACKGROUND)
    sim.init_particles(0, half, (0.05,) * 3, seed=seed)
    sim.init_particles(1, ppc - half, (0.006,) * 3, seed=seed)
    rng = np.random.default_rng(seed)
    nc = sim.ncell
    sim.upload_fields(rng.normal(0.0, sigma_E, size=(nc, 3)), np.tile(np.asarray(B0, dtype=np.float64), (nc, 1)))
    return sim


def generate(n, seeds, pairs_per_seed, ppc=64, field_substeps=8, burn_in=0, log=None):
    """Runs the simulator once per seed and records a pair at every step (after burn_in)."""
    snaps = Snapshots(n, len(seeds) * pairs_per_seed)
    V = snaps.V
    k = 0
    for seed in seeds:
        t0 = time.perf_counter()
        sim = config2_simulation(n, ppc, seed, field_substeps=field_substeps)
        for step in range(burn_in + pairs_per_seed):
            rec = step >= burn_in
            if rec:
                fs = sim.field_state()
                p = snaps.ptrs(k)
                dev_copy(p["E"], fs["E"], 16 * V)
                dev_copy(p["Bc"], fs["Bc"], 16 * V)
                dev_copy(p["J"], fs["J"], 16 * V)
            sim.field_step(use_operator=False)
            if rec:
                fs = sim.field_state()
                _chk(nop_lib().ipic_dev_sub4(p["tE"], fs["E"], p["E"], V, None))
                _chk(nop_lib().ipic_dev_sub4(p["tB"], fs["Bc"], p["Bc"], V, None))
                snaps.meta.append((seed, step))
                k += 1
            sim.particle_step()
        sim.close()
        if log:
            log(f"seed {seed}: {pairs_per_seed} pairs in {time.perf_counter() - t0:.2f} s")
    _chk(nop_lib().ipic_dev_sync(None))
    snaps.filled = k
    return snaps


def fit_scales(snaps, dt_field, samples=8):
    """Input/output scales and loss variances from the RMS of a few recorded pairs.  The
    potential's scale is the RMS of dt_field * E (the leapfrog B update is -curl+(dt E))."""
    idx = np.linspace(0, snaps.filled - 1, min(samples, snaps.filled)).astype(int)
    acc = {k: 0.0 for k in ("E", "Bc", "J", "tE", "tB")}
    for i in idx:
        d = snaps.download(int(i))
        for k in acc:
            acc[k] += float(np.mean(d[k].astype(np.float64) ** 2))
    rms = {k: np.sqrt(v / len(idx)) for k, v in acc.items()}
    tiny = 1e-30
    scales = {"in_E": max(rms["E"], tiny), "in_B": max(rms["Bc"], rms["E"], tiny), "in_J": max(rms["J"], tiny),
              "out_E": max(rms["tE"], tiny), "out_P": max(dt_field * rms["E"], tiny)}
    return scales, max(rms["tE"] ** 2, tiny), max(rms["tB"] ** 2, tiny)


def pack_all(net, snaps):
    """Packed BF16 inputs of every pair (V x 16 uint16 each)."""
    u = Dev(2 * 16 * snaps.V * snaps.filled)
    for i in range(snaps.filled):
        p = snaps.ptrs(i)
        net.pack(p["E"], p["Bc"], p["J"], u.ptr + 2 * 16 * snaps.V * i)
    net.sync()
    return u


def train(net, snaps, u_all, train_idx, varE, varB, epochs=1, lr=1e-3, lr_end=None, seed=0, log=None,
          log_every=200, k0=0, total=None):
    """Adam over the pairs of train_idx, one pair per step, shuffled each epoch; cosine decay
    from lr to lr_end over `total` steps (default: this call's), of which this call runs steps
    k0, k0 + 1, ... (a schedule spread over several calls on chunks of pairs).  Returns the
    per-step losses."""
    rng = np.random.default_rng(seed)
    lr_end = lr * 0.05 if lr_end is None else lr_end
    total = epochs * len(train_idx) if total is None else int(total)
    losses = []
    k = int(k0)
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(train_idx)
        for i in order:
            p = snaps.ptrs(int(i))
            l = net.loss_and_grad(p["tE"], p["tB"], varE, varB, u_ptr=u_all.ptr + 2 * 16 * snaps.V * int(i))
            a = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * k / max(total - 1, 1)))
            net.adam(a)
            losses.append(l)
            k += 1
            if log and (k - k0) % log_every == 0:
                log(f"step {k}/{total} epoch {ep} loss {np.mean(losses[-log_every:]):.4e} "
                    f"({time.perf_counter() - t0:.1f} s)")
    net.sync()
    return losses


def evaluate(net, snaps, u_all, idx, cell_width=(1.0, 1.0, 1.0), B0=(0.0, 0.2, 1.0), E0=(0.0, 0.0, 0.0)):
    """Errors of the operator against leapfrog on the pairs of idx (host-side, FP64):
    rel_total = |F_op - F_lf| / |F_lf| over E, B with the uniform background included;
    rel_increment = |dF_op - dF_lf| / |dF_lf|; the same for the persistence baseline
    (dF = 0) is 1 by definition and its total error is reported for comparison."""
    V = snaps.V
    psi = Dev(16 * V)
    E = Dev(16 * V)
    B = Dev(16 * V)
    res = []
    for i in idx:
        p = snaps.ptrs(int(i))
        dev_copy(E.ptr, p["E"], 16 * V)
        dev_copy(B.ptr, p["Bc"], 16 * V)
        net.forward(u_all.ptr + 2 * 16 * V * int(i), train=False)
        net.apply(E.ptr, B.ptr, psi.ptr, cell_width)
        net.sync()
        d = snaps.download(int(i))
        e_op = E.download(np.float32).reshape(V, 4)[:, :3].astype(np.float64)
        b_op = B.download(np.float32).reshape(V, 4)[:, :3].astype(np.float64)
        e0, b0 = d["E"].astype(np.float64), d["Bc"].astype(np.float64)
        e_lf, b_lf = e0 + d["tE"], b0 + d["tB"]
        diff2 = np.sum((e_op - e_lf) ** 2) + np.sum((b_op - b_lf) ** 2)
        tot2 = np.sum((e_lf + np.asarray(E0)) ** 2) + np.sum((b_lf + np.asarray(B0)) ** 2)
        inc2 = np.sum(d["tE"].astype(np.float64) ** 2) + np.sum(d["tB"].astype(np.float64) ** 2)
        evo2 = np.sum(e_lf ** 2) + np.sum(b_lf ** 2)  # the evolving field: no uniform background
        res.append({"rel_total": np.sqrt(diff2 / tot2), "rel_increment": np.sqrt(diff2 / max(inc2, 1e-300)),
                    "rel_evolving": np.sqrt(diff2 / max(evo2, 1e-300)),
                    "persistence_rel_total": np.sqrt(inc2 / tot2),
                    "persistence_rel_evolving": np.sqrt(inc2 / max(evo2, 1e-300))})
    psi.free(); E.free(); B.free()
    return {k: float(np.mean([r[k] for r in res])) for k in res[0]} | \
        {"max_" + k: float(np.max([r[k] for r in res])) for k in ("rel_total", "rel_increment", "rel_evolving")}


def forward_passes(n, reps, seed=0):
    """`reps` inference passes of an untrained operator on an n^3 grid of random fields (a fixed
    workload for kernel-level profiling of the convolutions)."""
    V = n ** 3
    scales = {"in_E": 1e-3, "in_B": 1e-3, "in_J": 1e-4, "out_E": 1e-4, "out_P": 5e-5}
    net = Network((n, n, n), scales, params=init_params(seed), train=False)
    rng = np.random.default_rng(seed)
    fields = []
    for _ in range(3):
        a = np.zeros((V, 4), dtype=np.float32)
        a[:, :3] = rng.normal(0.0, 1e-3, size=(V, 3))
        d = Dev(16 * V)
        d.upload(a)
        fields.append(d)
    net.pack(fields[0].ptr, fields[1].ptr, fields[2].ptr)
    for _ in range(int(reps)):
        net.forward()
    net.sync()


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="inference passes of the operator on random fields (profiling)")
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    forward_passes(a.n, a.reps)
    print(f"{a.reps} forward passes at {a.n}^3", flush=True)
