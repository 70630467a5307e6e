"""CPU-side checks of the drop-in boundary: the C-ABI library loads (no GPU needed for
dlopen) and exports every entry point include/ipic_gpu.h declares; the host package raises
instead of falling back when the library is absent."""
import ctypes
import os
import re

import pytest

import allscale_ipic3d_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|double|const char\*)\s+(ipic_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_binding_symbols():
    assert _declared("ipic_gpu.h") == sorted(G.ABI_SYMBOLS)


@pytest.mark.skipif(not os.path.exists(G.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    for name in _declared("ipic_gpu.h"):
        assert hasattr(lib, name), name


def test_operator_header_declares_binding_symbols():
    from allscale_ipic3d_amd import field_operator as FO
    assert _declared("ipic_nop.h") == sorted(FO.NOP_SYMBOLS)


@pytest.mark.skipif(not os.path.exists(G.LIB_PATH), reason="library not built")
def test_library_exports_every_operator_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    for name in _declared("ipic_nop.h"):
        assert hasattr(lib, name), name


def test_flag_constants_match_header():
    # every IPIC_FLAG_* of the header has the same value in the Python binding, and the bits
    # are distinct
    txt = open(os.path.join(ROOT, "include", "ipic_gpu.h")).read()
    flags = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define IPIC_(FLAG_\w+)\s+(\d+)u", txt)}
    assert "FLAG_REPRODUCIBLE" in flags and len(flags) >= 7
    for name, value in flags.items():
        assert getattr(G, name) == value, name
    values = list(flags.values())
    assert all(v & (v - 1) == 0 for v in values) and len(set(values)) == len(values)


def test_missing_library_raises(tmp_path):
    saved = G._lib
    G._lib = None
    try:
        with pytest.raises(G.IpicError):
            G.load_library(str(tmp_path / "nope.so"))
    finally:
        G._lib = saved


def test_desc_layout_matches_header():
    # ipic_desc: 4 int32, 3+3 doubles, 2 doubles, 8 species x 2 doubles, 2 int32, 3 doubles,
    # 1 double, 3 int32, 1 uint32, 1 double
    assert ctypes.sizeof(G.Desc) == 16 + 48 + 16 + 128 + 8 + 24 + 8 + 16 + 8
    assert ctypes.sizeof(G.StepStats) == 17 * 8
    # ipic_field_operator: pointer, 5 floats, 2 int32, 3 doubles (4 bytes of padding)
    assert ctypes.sizeof(G.FieldOperatorDesc) == 8 + 20 + 8 + 4 + 24
    assert ctypes.sizeof(G.Conservation) == 9 * 8


def _create(**kw):
    """ipic_gpu_create with a valid 4^3 desc modified by kw; returns (rc, last_error)."""
    import ctypes as C
    L = G.load_library()
    d = G.Desc()
    d.n[:] = kw.get("n", (4, 4, 4))
    d.cell_width[:] = kw.get("w", (1.0, 1.0, 1.0))
    d.dt, d.c = kw.get("dt", 0.1), kw.get("c", 1.0)
    d.nspecies = kw.get("nspecies", 1)
    d.species[0].q, d.species[0].qom = -1.0, -1.0
    d.field_substeps = kw.get("field_substeps", 1)
    d.tile[:] = kw.get("tile", (0, 0, 0))
    h = C.c_void_p()
    rc = L.ipic_gpu_create(C.byref(d), C.byref(h))
    if rc == 0:
        L.ipic_gpu_destroy(h)
    return rc, L.ipic_gpu_last_error(None).decode()


@pytest.mark.parametrize("kw,code,msg", [
    ({"n": (0, 4, 4)}, -1, "grid size"),
    ({"w": (1.0, -1.0, 1.0)}, -1, "cell width"),
    ({"dt": 0.0}, -1, "dt and c"),
    ({"c": -1.0}, -1, "dt and c"),
    ({"nspecies": 0}, -1, "nspecies"),
    ({"nspecies": 9}, -1, "nspecies"),
    ({"n": (2048, 1024, 1024)}, -4, "grid too large"),
    ({"n": (1, 1, 2097153)}, -4, "cells per axis"),
    ({"field_substeps": -1}, -1, "field_substeps"),
    ({"tile": (3, 0, 0)}, -1, "tile size"),
    ({"tile": (16, 0, 0), "n": (32, 4, 4)}, -1, "tile size"),
])
def test_create_rejects_invalid_descriptions(kw, code, msg):
    """Argument checks of ipic_gpu_create run before any device call (no GPU needed)."""
    rc, err = _create(**kw)
    assert rc == code and msg in err, (rc, err)
