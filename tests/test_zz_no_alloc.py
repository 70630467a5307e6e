"""ipic_nop_apply_noalloc (the operator head without the psi4 workspace, its potential in the
library's file-scope device array) against ipic_nop_apply on the same inputs, through the C
ABI: two calls in a row, the second on a smaller grid, each bitwise equal to the workspace
entry point (the same kernels on the same values); grids above the array are refused before
any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(FO, n, seed):
    rng = np.random.default_rng(seed)
    V = n[0] * n[1] * n[2]
    out = np.zeros((V, 16), dtype=np.float32)
    out[:, :6] = rng.normal(size=(V, 6)).astype(np.float32)
    dE = np.zeros((V, 4), dtype=np.float32)
    dB = np.zeros((V, 4), dtype=np.float32)
    dE[:, :3] = rng.normal(size=(V, 3))
    dB[:, :3] = rng.normal(size=(V, 3))
    return out, dE, dB


def _dev(FO, a):
    d = FO.Dev(a.nbytes)
    d.upload(np.ascontiguousarray(a))
    return d


def test_noalloc_apply_matches_workspace_apply_twice():
    from allscale_ipic3d_amd import field_operator as FO
    L = FO.nop_lib()
    sc = FO.Scales(1.0, 1.0, 1.0, 0.7, 0.3)
    iw = (C.c_float * 3)(1.0, 0.5, 2.0)
    for n, seed in (((16, 16, 32), 1), ((8, 4, 16), 2)):  # the second call on a smaller grid
        V = n[0] * n[1] * n[2]
        out, dE, dB = _case(FO, n, seed)
        n3 = (C.c_int32 * 3)(*n)
        o = _dev(FO, out)
        ea, ba, psi = _dev(FO, dE), _dev(FO, dB), FO.Dev(V * 16)
        eb, bb = _dev(FO, dE), _dev(FO, dB)
        FO._chk(L.ipic_nop_apply(n3, o.ptr, C.byref(sc), iw, ea.ptr, ba.ptr, psi.ptr, None))
        FO._chk(L.ipic_nop_apply_noalloc(n3, o.ptr, C.byref(sc), iw, eb.ptr, bb.ptr, None))
        FO._chk(L.ipic_dev_sync(None))
        Ea, Ba = ea.download(np.float32, V * 4), ba.download(np.float32, V * 4)
        Eb, Bb = eb.download(np.float32, V * 4), bb.download(np.float32, V * 4)
        print(n, "max |dE| change", float(np.abs(Ea.reshape(-1, 4) - dE).max()),
              "max |dBc| change", float(np.abs(Ba.reshape(-1, 4) - dB).max()), flush=True)
        assert np.abs(Ba.reshape(-1, 4) - dB).max() > 0  # the head changed the state
        np.testing.assert_array_equal(Eb, Ea)
        np.testing.assert_array_equal(Bb, Ba)
        for d in (o, ea, ba, psi, eb, bb):
            d.free()


def test_noalloc_apply_refuses_grids_above_its_array():
    from allscale_ipic3d_amd import field_operator as FO
    L = FO.nop_lib()
    sc = FO.Scales(1.0, 1.0, 1.0, 0.7, 0.3)
    n3 = (C.c_int32 * 3)(256, 128, 128)  # 2^22 voxels: refused before any pointer is read
    one = FO.Dev(256)
    rc = L.ipic_nop_apply_noalloc(n3, one.ptr, C.byref(sc), None, one.ptr, one.ptr, None)
    assert rc == -4 and b"IPIC_NOP_NOALLOC_MAX_VOXELS" in L.ipic_nop_last_error()
    one.free()
