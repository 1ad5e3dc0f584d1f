"""C ABI checks that need no GPU: the library loads, exports every entry
point include/*.h declares, and its pure host helpers follow the reference's
width rules. No kernel is launched here."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(am_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def L():
    import amethyst_amd
    return amethyst_amd.lib()


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("am_filter", "am_compressstore", "am_binary", "am_sum", "am_groupby", "am_convert"):
        assert must in names


def test_library_exports_every_declared_symbol(L):
    path = os.path.join(ROOT, "amethyst_amd", "libamethyst.so")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    """The device code object bundled into the library targets gfx950."""
    path = os.path.join(ROOT, "amethyst_amd", "libamethyst.so")
    blob = open(path, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_packed_words(L):
    import amethyst_amd as am
    n = 1000
    assert am.packed_words(am.BASELINE, 11, n) == n
    assert am.packed_words(am.SLAB, 11, n) == -(-n // 5)
    assert am.packed_words(am.BITSTREAM, 11, n) == -(-(n * 11) // 64)
    assert am.packed_words(am.U32, 11, n) == n // 2
    assert am.packed_words(am.TIGHT, 12, n) == -(-(n * 2) // 8)
    assert am.packed_words(am.U32, 33, n) == 0  # invalid combination
    assert am.packed_words(am.SLAB, 0, n) == 0


def test_outbits_and_pred_width(L):
    import amethyst_amd as am
    # op_add_outbits / op_mul_outbits / op_mod_outbits (src/config.cuh:44-55)
    assert [am.outbits(am.ADD, w) for w in (1, 11, 63, 64)] == [2, 12, 64, 64]
    assert [am.outbits(am.MUL, w) for w in (1, 11, 32, 40)] == [2, 22, 64, 64]
    assert [am.outbits(am.MOD, w) for w in (1, 11, 64)] == [1, 11, 64]
    assert am.pred_width(am.LT, 1024, 11) == 10
    assert am.pred_width(am.LE, 1024, 11) == 11
    assert am.pred_width(am.GE, 1024, 11) == 11
    assert am.pred_width(am.LT, 1, 11) == 1
    assert am.pred_width(am.EQ, 5, 11) == 3
    assert L.am_gen_bitmask(64) == 2**64 - 1
    assert L.am_gen_bitmask(11) == 2047


def test_status_strings_and_errors(L):
    import amethyst_amd as am
    assert L.am_status_string(0) == b"ok"
    assert L.am_status_string(2) == b"unsupported configuration"
    for code in range(8):  # every am_status has its own message
        assert L.am_status_string(code) != b"unknown status", code
    assert L.am_status_string(8) == b"unknown status"
    assert L.am_workspace_status(None) == 1  # null workspace: invalid, checked before any HIP call
    # invalid arguments are rejected on the host before any launch
    bad = am.am_column(None, 10, 11, am.SLAB)
    ok = am.am_column(None, 0, 11, am.SLAB)
    assert L.am_convert(ctypes.byref(bad), ctypes.byref(ok), None) == 1
    col65 = am.am_column(None, 0, 65, am.SLAB)
    assert L.am_convert(ctypes.byref(col65), ctypes.byref(ok), None) == 1
    u33 = am.am_column(None, 0, 33, am.U32)
    assert L.am_filter(ctypes.byref(u33), am.am_predicate(0, 0, 0), None, None) == 1
    keys25 = am.am_column(ctypes.c_void_p(8), 1, 25, am.SLAB)
    keys20 = am.am_column(ctypes.c_void_p(8), 1, 20, am.SLAB)
    vals = am.am_column(ctypes.c_void_p(8), 1, 20, am.SLAB)
    dummy = ctypes.c_void_p(8)
    # keys wider than 24 bits: unsupported
    assert L.am_groupby(ctypes.byref(keys25), ctypes.byref(vals), None, 0, dummy, dummy, dummy, dummy, dummy,
                        1 << 30, None) == 2
    # 13..24-bit keys need the partitioned path's workspace (checked before any launch)
    need = L.am_groupby_workspace_bytes(1, 20, 20)
    assert need > L.am_groupby_workspace_bytes(1, 12, 20) > 0
    assert L.am_groupby(ctypes.byref(keys20), ctypes.byref(vals), None, 0, dummy, dummy, dummy, dummy, dummy,
                        need - 1, None) == 4
    assert L.am_groupby_workspace_bytes(1, 25, 20) == 0 and L.am_groupby_workspace_bytes(1, 10, 33) == 0
    # missing workspace
    assert L.am_sum(ctypes.byref(vals), dummy, None, 0, None) == 4


def test_tune_table_parsing(L, tmp_path):
    """The launch-shape table's text format and bounds (host only)."""
    import amethyst_amd as am
    saved = [(k, b, fi, fo, L.am_tune_get(k, b, fi, fo)) for k in range(6) for b in range(1, 65)
             for fi in range(5) for fo in range(5)]
    saved = [x for x in saved if x[4]]
    try:
        L.am_tune_clear()
        assert L.am_tune_get(am.KNOB_FILTER_WORDS, 11, am.SLAB, 0) == 0
        good = tmp_path / "t.txt"
        good.write_text("# comment\n\n0 11 1 0 4   # filter\n1 11 1 4 6\n3 17 2 0 1024\n")
        assert L.am_tune_load(str(good).encode()) == 3
        assert L.am_tune_get(0, 11, 1, 0) == 4 and L.am_tune_get(1, 11, 1, 4) == 6 and L.am_tune_get(3, 17, 2, 0) == 1024
        assert L.am_tune_set(0, 11, 1, 0, 0) == 0 and L.am_tune_get(0, 11, 1, 0) == 0
        assert L.am_tune_set(6, 11, 1, 0, 1) == 1      # unknown knob
        assert L.am_tune_set(0, 65, 1, 0, 1) == 1      # width
        assert L.am_tune_set(0, 11, 5, 0, 1) == 1      # format
        bad = tmp_path / "b.txt"
        bad.write_text("0 11 1 0 4\n0 11 1\n")
        assert L.am_tune_load(str(bad).encode()) == -1
        assert L.am_tune_load(str(tmp_path / "missing.txt").encode()) == -1
        # the committed table, if present, parses
        if os.path.exists(am.TUNE_TABLE):
            assert L.am_tune_load(am.TUNE_TABLE.encode()) >= 0
    finally:
        L.am_tune_clear()
        for k, b, fi, fo, v in saved:
            L.am_tune_set(k, b, fi, fo, v)
