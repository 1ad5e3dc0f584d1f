"""am_compressstore_noalloc (fixed per-GPU scratch, no allocation) against
am_compressstore (caller's workspace): identical counts, masks and packed
words, on calls in a row that share the scratch -- each later call smaller
than the one before, so stale descriptors of a longer call are present."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(am, col, out, stats, mask=None, pred=None, mask_out=None):
    L = am.lib()
    a, o = col.c(), out.c()
    pp = ctypes.byref(am.am_predicate(pred[0], 0, pred[1])) if pred is not None else None
    mp = ctypes.c_void_p(mask.data_ptr()) if mask is not None else None
    mop = ctypes.c_void_p(mask_out.data_ptr()) if mask_out is not None else None
    return L.am_compressstore_noalloc(ctypes.byref(a), mp, pp, ctypes.byref(o), mop,
                                      ctypes.c_void_p(stats.data_ptr()), am._stream())


def _compare(am, col, out_bits, out_fmt, pred=None, mask=None):
    import torch
    n = col.n
    ref_out, ref_stats, ref_mask = am.compressstore(col, out_bits, out_fmt, pred=pred, mask=mask, want_mask=True)
    out = am.Column(n, out_bits, out_fmt)
    stats = am.device_words(2)
    mo = am.device_words((n + 63) // 64)
    st = _noalloc(am, col, out, stats, mask=mask, pred=pred, mask_out=mo)
    assert st == 0, am.STATUS.get(st, st)
    torch.cuda.synchronize()
    rs, gs = am.u64(ref_stats), am.u64(stats)
    assert list(map(int, gs)) == list(map(int, rs)), (list(gs), list(rs))
    k = min(int(rs[0]), n)
    used = am.packed_words(out_fmt, out_bits, k)
    assert np.array_equal(out.host_words()[:used], ref_out.host_words()[:used])
    nmw = (n + 63) // 64
    assert torch.equal(mo[:nmw], ref_mask[:nmw])
    return int(rs[0])


def test_noalloc_matches_workspace_entry_point(am, oracle_mod):
    import torch
    cases = [
        # (n, bits, input format, output bits, output format, predicate or mask selectivity)
        ((1 << 24) + 77, 11, am.SLAB, 32, am.U32, (am.LT, 1024)),     # headline shape, container out
        (3_000_001, 11, am.SLAB, 11, am.SLAB, (am.LT, 700)),           # packed out (window look-back)
        (200_003, 13, am.BITSTREAM, 10, am.BITSTREAM, (am.GE, 4000)),   # element kernel, overflow
        (5003, 4, am.SLAB, 4, am.SLAB, 0.5),                            # mask-driven
        (1, 7, am.SLAB, 7, am.U32, (am.LT, 100)),
    ]
    for n, bits, fi, ob, fo, sel in cases:
        print("noalloc", n, bits, fi, ob, fo, sel, flush=True)
        col = am.gen_uniform(n + bits, n, bits, fi)
        if isinstance(sel, tuple):
            k = _compare(am, col, ob, fo, pred=sel)
        else:
            m = am.gen_mask_bernoulli(3, n, sel)
            k = _compare(am, col, ob, fo, mask=m)
        # the oracle agrees on the count
        words = col.host_words()
        if isinstance(sel, tuple):
            exp = oracle_mod.filter(fi, bits, words, n, sel[0], sel[1])
            assert k == int(sum(bin(int(x)).count("1") for x in exp))
        del col
        torch.cuda.empty_cache()


def test_noalloc_rejects_oversized_input(am):
    n = (1 << 28) + 1
    col = am.Column(0, 11, am.SLAB)
    big = am.am_column(col.buf.data_ptr(), n, 11, am.SLAB)  # never launched: rejected on the host
    out = am.am_column(col.buf.data_ptr(), n, 32, am.U32)
    stats = am.device_words(2)
    L = am.lib()
    st = L.am_compressstore_noalloc(ctypes.byref(big), None, ctypes.byref(am.am_predicate(am.LT, 0, 5)),
                                    ctypes.byref(out), None, ctypes.c_void_p(stats.data_ptr()), am._stream())
    assert st == 7
