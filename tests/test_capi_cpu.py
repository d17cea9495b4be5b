"""The C ABI library loads without a GPU, exports exactly what include/compas_amd.h declares, its
struct sizes match the ctypes mirrors, and its host-only helpers and argument checks work (no
call here reaches a kernel)."""
import ctypes
import os
import re

import pytest

from compas_amd import Layout, capi, make_geometry, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "compas_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cmp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_abi_struct_sizes_match():
    lib = capi.load()
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.cmp_abi_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert a.value == ctypes.sizeof(capi.cmp_params)
    assert b.value == ctypes.sizeof(capi.cmp_geometry)
    assert c.value == ctypes.sizeof(capi.cmp_box)
    assert b"gfx950" in lib.cmp_version()


def test_level_layout_offsets():
    lay = Layout(2, [([0, 0], [15, 7]), ([16, 0], [31, 7])], 3, 6)
    assert lay.cboxes[0].offset == 0
    assert lay.cboxes[1].offset == 22 * 14 * 6
    assert lay.total == 2 * 22 * 14 * 6
    lay3 = Layout(3, [([0, 0, 0], [7, 7, 7])], 3, 7)
    assert lay3.total == 14 ** 3 * 7
    lib = capi.load()
    b = capi.cmp_box()
    b.lo[0], b.lo[1], b.hi[0], b.hi[1] = 0, 0, 9, 4
    assert lib.cmp_fab_size(2, ctypes.byref(b), 2, 3) == 14 * 9 * 3


def test_argument_errors_are_returned_not_raised():
    """Bad arguments come back as status codes before any GPU work (the reference aborts)."""
    lib = capi.load()
    P = make_params(2)
    G = make_geometry((16, 16), (0, 0), (1, 1), (1, 1))
    b = capi.cmp_box()
    b.hi[0], b.hi[1] = 15, 15
    # null state pointer
    assert lib.cmp_rhs_box(ctypes.byref(P), ctypes.byref(G), ctypes.byref(b), 3, None, None, None,
                           None, None, None, None, None, None, None) == capi.CMP_ERR_ARG
    # too few ghost cells for WENO5
    assert lib.cmp_rhs_box(ctypes.byref(P), ctypes.byref(G), ctypes.byref(b), 2, ctypes.c_void_p(8),
                           ctypes.c_void_p(8), None, None, None, None, None, None, None,
                           None) == capi.CMP_ERR_ARG
    # bound-preserving schemes are not built here
    Pb = make_params(2)
    Pb.id_bound = 1
    assert lib.cmp_rhs_box(ctypes.byref(Pb), ctypes.byref(G), ctypes.byref(b), 3, ctypes.c_void_p(8),
                           ctypes.c_void_p(8), None, None, None, None, None, None, None,
                           None) == capi.CMP_ERR_UNSUPPORTED
    # LLF is not on this path
    Pl = make_params(2)
    Pl.id_riemann = 0
    assert lib.cmp_rhs_box(ctypes.byref(Pl), ctypes.byref(G), ctypes.byref(b), 3, ctypes.c_void_p(8),
                           ctypes.c_void_p(8), None, None, None, None, None, None, None,
                           None) == capi.CMP_ERR_UNSUPPORTED
    # dimension mismatch between params and geometry
    G3 = make_geometry((16, 16, 16), (0, 0, 0), (1, 1, 1), (1, 1, 1))
    assert lib.cmp_rhs_box(ctypes.byref(P), ctypes.byref(G3), ctypes.byref(b), 3, ctypes.c_void_p(8),
                           ctypes.c_void_p(8), None, None, None, None, None, None, None,
                           None) == capi.CMP_ERR_ARG
    h = ctypes.c_void_p()
    assert lib.cmp_level_create(ctypes.byref(P), ctypes.byref(G), 0, None, 3,
                                ctypes.byref(h)) == capi.CMP_ERR_ARG
    assert lib.cmp_level_destroy(None) == 0
    assert lib.cmp_level_check(None, None) == capi.CMP_ERR_ARG


def test_tile_cache_file_round_trip(tmp_path):
    """The tile autotuner's cache as a text file (host code only): a file of two swept box shapes
    loads, saves back to the same entries, and a truncated file is refused with CMP_ERR_ARG."""
    lib = capi.load()
    src = tmp_path / "in.txt"
    src.write_text("2 5 5 2 0 3 3 16 16 1 0.47 2  16 15 1 0.471 0.52  16 16 1 0.444 0.46\n"
                   "3 5 5 2 1 0 3 32 32 32 0.2076 1  8 8 3 0.21 0.25\n")
    assert lib.cmp_tile_cache_clear() == 0
    assert lib.cmp_tile_cache_load(str(src).encode()) == 0
    out = tmp_path / "out.txt"
    assert lib.cmp_tile_cache_save(str(out).encode()) == 0

    def parse(p):
        return sorted(tuple(float(v) for v in line.split()) for line in p.read_text().splitlines() if line.strip())

    assert parse(out) == parse(src)
    bad = tmp_path / "bad.txt"
    bad.write_text("2 5 5 2 0 3 3 16 16 1 0.47 2  16 15 1 0.471\n")
    assert lib.cmp_tile_cache_load(str(bad).encode()) == capi.CMP_ERR_ARG
    assert lib.cmp_tile_cache_load(str(tmp_path / "missing.txt").encode()) == capi.CMP_ERR_ARG
    assert lib.cmp_tile_cache_clear() == 0
