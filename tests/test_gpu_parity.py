"""GPU parity: every operator through the C ABI vs the CPU oracle, bit-exact.

Inputs are seeded numpy draws packed by the oracle; outputs are compared as
packed words (decoded values AND dead/padding bits), counts and masks.
Edge cases follow the reference's validation loops: ragged lengths around
the 64-element mask word and the 4096-element tile, empty and full
selections, every width class (1, sub-byte, byte, straddling, 32, >32, 64).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMATS = [0, 1, 2, 3, 4]  # baseline, slab, bitstream, tight, u32
WIDTHS = [1, 3, 4, 8, 11, 17, 24, 31, 32, 33, 47, 63, 64]
SIZES = [1, 63, 64, 65, 4095, 4096, 4097, 20011]


def fmt_ok(fmt, bits):
    return not (fmt == 4 and bits > 32)


def rand_vals(rng, n, bits):
    hi = rng.integers(0, 2**32, size=n, dtype=np.uint64)
    lo = rng.integers(0, 2**32, size=n, dtype=np.uint64)
    v = (hi << np.uint64(32)) | lo
    m = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    return v & m


def upload(am, oracle, vals, bits, fmt):
    words = oracle.pack(fmt, bits, vals)
    return am.Column.from_host_words(words, len(vals), bits, fmt)


def say(*a):
    print(*a, flush=True)


def test_prng_matches_reference_stream(am, oracle_mod):
    for seed in (42, 0, 7):
        for off in (0, 1, 1000):
            say("prng", seed, off)
            got = am.prng_fill(seed, 5000, offset=off).cpu().numpy().view(np.uint32)
            exp = oracle_mod.prng(seed, 5000, skip=off)
            assert np.array_equal(got, exp)


def test_stream_copy(am, oracle_mod):
    """The harness copy moves every word (odd lengths: the tail word too)."""
    import torch
    for n in (1, 2, 3, 1001, 4096, (1 << 20) + 1):
        say("stream_copy", n)
        src = torch.randint(-2**62, 2**62, (n,), dtype=torch.int64, device="cuda")
        dst = torch.zeros_like(src)
        am.stream_copy(src, dst)
        torch.cuda.synchronize()
        assert torch.equal(src, dst), n


def test_gen_uniform(am, oracle_mod):
    for fmt in FORMATS:
        for bits in (3, 11, 32, 40, 64):
            if not fmt_ok(fmt, bits):
                continue
            n = 3001
            say("gen_uniform", fmt, bits)
            col = am.gen_uniform(42, n, bits, fmt)
            if bits <= 32:
                vals = oracle_mod.prng(42, n).astype(np.uint64) & np.uint64((1 << bits) - 1)
            else:
                d = oracle_mod.prng(42, 2 * n).astype(np.uint64)
                v = (d[0::2] << np.uint64(32)) | d[1::2]
                vals = v & (np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF))
            assert np.array_equal(col.host_words(), oracle_mod.pack(fmt, bits, vals))


def test_convert_all_formats(am, oracle_mod):
    rng = np.random.default_rng(1)
    for bits in WIDTHS:
        vals = rand_vals(rng, 5003, bits)
        for fi in FORMATS:
            if not fmt_ok(fi, bits):
                continue
            col = upload(am, oracle_mod, vals, bits, fi)
            for fo in FORMATS:
                for ob in sorted({bits, max(1, bits - 2), min(64, bits + 5)}):
                    if not fmt_ok(fo, ob):
                        continue
                    say("convert", bits, fi, fo, ob)
                    out = am.convert(col, ob, fo)
                    exp = oracle_mod.pack(fo, ob, vals & np.uint64(oracle_mod.bitmask(ob)))
                    assert np.array_equal(out.host_words(), exp), (bits, fi, fo, ob)


def test_filter(am, oracle_mod):
    rng = np.random.default_rng(2)
    for bits in WIDTHS:
        for n in (1, 64, 4097, 20011):
            vals = rand_vals(rng, n, bits)
            for fmt in (0, 1, 2, 4):
                if not fmt_ok(fmt, bits):
                    continue
                col = upload(am, oracle_mod, vals, bits, fmt)
                half = 1 << (bits - 1)
                for op, c in ((3, half), (0, half), (4, int(vals[0])), (5, int(vals[0])), (1, 0), (2, 2**64 - 1)):
                    say("filter", bits, n, fmt, op, c)
                    got = am.u64(am.filter(col, op, c))[:(n + 63) // 64]
                    exp = oracle_mod.filter(fmt, bits, oracle_mod.pack(fmt, bits, vals), n, op, c)
                    assert np.array_equal(got, exp), (bits, n, fmt, op)


def test_filter_pow2_fields_every_threshold(am, oracle_mod):
    """2-, 4- and 8-bit slabs, whose fields are selected SWAR-wise under
    ranges from 0 or up to the top value: every threshold of <, <=, >, >=,
    and == / != of 0 and of the top value, in the filter and in the fused
    compaction."""
    rng = np.random.default_rng(29)
    for bits in (2, 4, 8):
        vmax = (1 << bits) - 1
        n = 20011
        vals = rand_vals(rng, n, bits)
        col = upload(am, oracle_mod, vals, bits, 1)
        cs = sorted({0, 1, vmax // 2, vmax - 1, vmax, vmax + 1} | (set(range(vmax + 2)) if bits <= 4 else set()))
        cases = [(op, c) for op in (0, 1, 2, 3) for c in cs] + [(4, 0), (4, vmax), (5, 0), (5, vmax)]
        for op, c in cases:
            say("filter-pow2", bits, op, c)
            got = am.u64(am.filter(col, op, c))[:(n + 63) // 64]
            exp = oracle_mod.filter(1, bits, oracle_mod.pack(1, bits, vals), n, op, c)
            assert np.array_equal(got, exp), (bits, op, c)
        for op, c in ((0, vmax // 2 + 1), (3, vmax // 2), (2, vmax - 1)):
            say("compress-pow2", bits, op, c)
            _check_compress(am, oracle_mod, vals, bits, 1, bits, 1, pred=(op, c))
            _check_compress(am, oracle_mod, vals, bits, 1, 32, 4, pred=(op, c))


def _check_compress(am, oracle_mod, vals, bits, fi, ob, fo, mask_words=None, pred=None, cap=None, ws=None):
    n = len(vals)
    col = upload(am, oracle_mod, vals, bits, fi)
    if pred is not None:
        mask_words = oracle_mod.filter(fi, bits, oracle_mod.pack(fi, bits, vals), n, pred[0], pred[1])
        out, stats, mo = am.compressstore(col, ob, fo, pred=pred, want_mask=True, capacity=cap, ws=ws)
    else:
        import torch
        m = torch.from_numpy(np.ascontiguousarray(mask_words).view(np.int64)).cuda()
        out, stats, mo = am.compressstore(col, ob, fo, mask=m, want_mask=True, capacity=cap, ws=ws)
    k, exp_words, ovf = oracle_mod.compressstore(fi, bits, oracle_mod.pack(fi, bits, vals), n, mask_words, fo, ob,
                                                 out_cap=cap)
    st = am.u64(stats)
    assert int(st[0]) == k
    assert int(st[1]) == ovf
    # the mask written back has the input's selection; bits past n are zero
    exp_mask = np.array(mask_words[:(n + 63) // 64], dtype=np.uint64)
    if n % 64:
        exp_mask[-1] &= np.uint64((1 << (n % 64)) - 1)
    assert np.array_equal(am.u64(mo)[:(n + 63) // 64], exp_mask)
    # words the oracle defines: those holding elements < min(k, cap)
    kk = min(k, out.n)
    used = oracle_mod.packed_words(fo, ob, kk)
    got = out.host_words()[:used]
    assert np.array_equal(got, exp_words[:used]), (bits, fi, ob, fo, n)


def test_compress_mask_formats(am, oracle_mod):
    rng = np.random.default_rng(3)
    ws = am.Workspace(1 << 16)
    for bits in (1, 4, 11, 17, 32, 33, 64):
        for n in SIZES:
            vals = rand_vals(rng, n, bits)
            for sel in (0.0, 0.01, 0.5, 1.0):
                mbits = rng.random(n) < sel
                mask = np.packbits(mbits, bitorder="little")
                mask = np.concatenate([mask, np.zeros((-len(mask)) % 8, np.uint8)]).view(np.uint64)
                for fi in (0, 1, 2, 3, 4):
                    if not fmt_ok(fi, bits):
                        continue
                    for fo, ob in ((1, bits), (2, bits), (0, bits), (4, min(bits, 32)), (1, max(1, bits - 1))):
                        if not fmt_ok(fo, ob):
                            continue
                        say("compress-mask", bits, n, sel, fi, fo, ob)
                        _check_compress(am, oracle_mod, vals, bits, fi, ob, fo, mask_words=mask, ws=ws)
    ws.check()


def test_compress_predicate_fused(am, oracle_mod):
    rng = np.random.default_rng(4)
    ws = am.Workspace(1 << 20)
    for bits in (4, 8, 11, 17, 24, 32):
        n = 300007
        vals = rand_vals(rng, n, bits)
        for frac in (0.01, 0.5, 0.99):
            c = int(round(frac * (1 << bits)))
            pw = am.pred_width(0, c, bits)
            for fo, ob in ((4, bits), (1, bits), (2, bits), (1, pw), (2, pw), (3, pw)):
                say("compress-pred", bits, frac, fo, ob)
                _check_compress(am, oracle_mod, vals, bits, 1, ob, fo, pred=(0, c), ws=ws)
    ws.check()


def test_compress_predicate_all_inputs(am, oracle_mod):
    """Predicate compaction from every input format and width class (word
    kernel for slab/containers, element kernel for bitstream), every
    predicate operator."""
    rng = np.random.default_rng(14)
    ws = am.Workspace(1 << 20)
    for bits in (1, 3, 5, 8, 9, 16, 21, 31, 32, 33, 40, 64):
        for n in (1, 1023, 1024, 1025, 65, 70001):
            vals = rand_vals(rng, n, bits)
            c = int(vals[n // 2])
            for fi in FORMATS:
                if not fmt_ok(fi, bits):
                    continue
                for op in range(6):
                    fo, ob = ((4, min(bits, 32)), (1, bits), (2, bits), (0, bits), (3, bits), (1, max(1, bits // 2)))[op]
                    say("compress-pred-in", bits, n, fi, op, fo, ob)
                    _check_compress(am, oracle_mod, vals, bits, fi, ob, fo, pred=(op, c), ws=ws)
    ws.check()


def test_compress_many_tiles(am, oracle_mod):
    """More than 64*64 tiles: the group look-back takes several steps."""
    rng = np.random.default_rng(15)
    n = (1 << 25) + 12345
    ws = am.Workspace(n)
    for bits, fi, fo, ob in ((11, 1, 4, 11), (11, 1, 1, 10), (30, 4, 2, 30), (7, 2, 0, 7)):
        vals = rand_vals(rng, n, bits)
        say("compress-many-tiles", bits, fi, fo, ob)
        _check_compress(am, oracle_mod, vals, bits, fi, ob, fo, pred=(0, 1 << (bits - 1)), ws=ws)
    ws.check()


def test_config_c1_exact(am, oracle_mod):
    """Config C1 as specified: N = 2^24 values rand() & 0x7FF from the PCG
    stream of seed 42, predicate v < 1024, output u32 + bitmask, slab and
    bitstream inputs -- generation, count, mask and output byte-exact with the
    oracle. Also the reference's own predicate v >= 2^(w-1) (the complement)."""
    n, w = 1 << 24, 11
    vals = oracle_mod.prng(42, n).astype(np.uint64) & np.uint64(0x7FF)
    ws = am.Workspace(n)
    for fmt in (1, 2):
        col = am.gen_uniform(42, n, w, fmt)
        packed = oracle_mod.pack(fmt, w, vals)
        assert np.array_equal(col.host_words(), packed), fmt
        for op, c in ((0, 1024), (3, 1 << (w - 1))):
            say("config-c1", fmt, op, c)
            mask = oracle_mod.filter(fmt, w, packed, n, op, c)
            out, stats, mo = am.compressstore(col, 32, 4, pred=(op, c), want_mask=True, ws=ws)
            k, exp_words, ovf = oracle_mod.compressstore(fmt, w, packed, n, mask, 4, 32)
            st = am.u64(stats)
            assert int(st[0]) == k and int(st[1]) == 0
            assert abs(k - n // 2) < 20000  # selectivity ~0.5
            assert np.array_equal(am.u64(mo)[:(n + 63) // 64], mask)
            used = oracle_mod.packed_words(4, 32, k)
            assert np.array_equal(out.host_words()[:used], exp_words[:used])
    ws.check()


def test_filter_full_size_matches_compress_mask(am, oracle_mod):
    """At 2^28 elements the word filter runs its widest tiles (8 words per
    lane, mask words gathered per wave), which the small parity cases do not
    reach. Size-independent check: its mask equals the mask the compaction
    kernel writes for the same predicate (an independent kernel, pinned to the
    oracle above), at the selectivity the predicate implies."""
    import torch
    n = 1 << 28
    ws = am.Workspace(n)
    for w in (3, 5, 7, 11, 17):
        col = am.gen_uniform(7 + w, n, w, 1)
        for op, c in ((0, 1 << (w - 1)), (3, (1 << w) // 3)):
            say("filter-full", w, op, c)
            fm = am.filter(col, op, c)
            out, stats, mo = am.compressstore(col, 32, 4, pred=(op, c), want_mask=True, ws=ws)
            torch.cuda.synchronize()
            nmw = (n + 63) // 64
            assert torch.equal(fm[:nmw], mo[:nmw]), (w, op, c)
            k = int(am.u64(stats)[0])
            frac = c / (1 << w) if op == 0 else ((1 << w) - c) / (1 << w)  # uniform w-bit values
            assert abs(k - n * frac) < n // 50, (w, op, c, k)
            del out, stats, mo, fm
    ws.check()


def test_compress_capacity_truncation(am, oracle_mod):
    rng = np.random.default_rng(5)
    ws = am.Workspace(1 << 16)
    vals = rand_vals(rng, 50000, 11)
    for cap in (0, 1, 100, 7777, 24000):
        for fo, ob in ((1, 11), (2, 11), (4, 11)):
            say("compress-cap", cap, fo)
            _check_compress(am, oracle_mod, vals, 11, 1, ob, fo, pred=(0, 1024), cap=cap, ws=ws)


def test_compress_workspace_reuse(am, oracle_mod):
    """Many launches on one workspace with varying tile counts (parity reset)."""
    rng = np.random.default_rng(6)
    ws = am.Workspace(1 << 18)
    for i, n in enumerate([200000, 1, 5000, 262144, 4096, 100, 150000, 9000]):
        vals = rand_vals(rng, n, 11)
        say("compress-reuse", i, n)
        _check_compress(am, oracle_mod, vals, 11, 1, 11, 2, pred=(0, 700), ws=ws)
    ws.check()


def test_binary_ops(am, oracle_mod):
    rng = np.random.default_rng(7)
    for bits in (1, 4, 8, 11, 17, 24, 31, 32, 40, 63, 64):
        n = 4099
        a = rand_vals(rng, n, bits)
        b = rand_vals(rng, n, bits)
        b[::7] = 0  # exercise div/mod by zero
        for fmt in (0, 1, 2, 3):
            ca = upload(am, oracle_mod, a, bits, fmt)
            cb = upload(am, oracle_mod, b, bits, fmt)
            for op in range(11):
                for ofmt in (fmt, 1 if fmt != 1 else 2):
                    ob = oracle_mod.outbits(op, bits)
                    say("binary", bits, fmt, op, ofmt, ob)
                    out = am.binary(op, ca, cb, out_bits=ob, out_fmt=ofmt)
                    exp = oracle_mod.binary(op, fmt, bits, oracle_mod.pack(fmt, bits, a), oracle_mod.pack(fmt, bits, b),
                                            n, ofmt, ob)
                    assert np.array_equal(out.host_words(), exp), (bits, fmt, op, ofmt)


def test_binary_bitstream_multiblock(am, oracle_mod):
    """Bitstream inputs over many blocks of the word-per-lane stream kernel
    (block-relative element and bit offsets, words straddling blocks), into
    bitstream and slab outputs."""
    rng = np.random.default_rng(17)
    for bits in (2, 3, 11, 17, 31, 32):
        n = 200003
        a = rand_vals(rng, n, bits)
        b = rand_vals(rng, n, bits)
        ca = upload(am, oracle_mod, a, bits, 2)
        cb = upload(am, oracle_mod, b, bits, 2)
        for op in (0, 1, 7, 9, 10):  # add, mul, and, xor, sub
            for ofmt in (2, 1):
                ob = oracle_mod.outbits(op, bits)
                say("binary-stream", bits, op, ofmt, ob)
                out = am.binary(op, ca, cb, out_bits=ob, out_fmt=ofmt)
                exp = oracle_mod.binary(op, 2, bits, oracle_mod.pack(2, bits, a), oracle_mod.pack(2, bits, b),
                                        n, ofmt, ob)
                assert np.array_equal(out.host_words(), exp), (bits, op, ofmt)


def test_binary_slab_multiblock(am, oracle_mod):
    """Slab inputs over many blocks into slab outputs of fewer fields per word
    (an output word's fields start mid-way through an input word and run into
    the next), and into containers."""
    rng = np.random.default_rng(27)
    for bits in (3, 5, 6, 7, 10, 13, 21):
        n = 300007
        a = rand_vals(rng, n, bits)
        b = rand_vals(rng, n, bits)
        ca = upload(am, oracle_mod, a, bits, 1)
        cb = upload(am, oracle_mod, b, bits, 1)
        for op in (0, 1, 10):  # add, mul, sub
            for ofmt in (1, 0):
                ob = oracle_mod.outbits(op, bits)
                say("binary-slab", bits, op, ofmt, ob)
                out = am.binary(op, ca, cb, out_bits=ob, out_fmt=ofmt)
                exp = oracle_mod.binary(op, 1, bits, oracle_mod.pack(1, bits, a), oracle_mod.pack(1, bits, b),
                                        n, ofmt, ob)
                assert np.array_equal(out.host_words(), exp), (bits, op, ofmt)


def test_sum(am, oracle_mod):
    rng = np.random.default_rng(8)
    ws = am.Workspace(0)
    for bits in (1, 2, 3, 4, 5, 8, 11, 16, 20, 21, 31, 32, 33, 64):
        for n in (0, 1, 4095, 4097, 1000003):
            vals = rand_vals(rng, n, bits)
            for fmt in (0, 1, 2, 3, 4):
                if not fmt_ok(fmt, bits):
                    continue
                say("sum", bits, n, fmt)
                col = upload(am, oracle_mod, vals, bits, fmt)
                got = int(am.u64(am.sum(col, ws=ws))[0])
                exp = oracle_mod.sum(fmt, bits, oracle_mod.pack(fmt, bits, vals), n)
                assert got == exp, (bits, n, fmt)


def test_sum_swar_saturated(am, oracle_mod):
    """Every field at its maximum, on columns long enough that each lane
    accumulates its slots up to the flush bound (K = 2^w words for slabs):
    the sum must be exactly n * (2^w - 1)."""
    ws = am.Workspace(0)
    for bits, n in ((2, 1 << 26), (3, 21 * (1 << 22)), (4, 1 << 27)):
        eps = 64 // bits
        words = (n + eps - 1) // eps
        word = np.uint64((1 << (eps * bits)) - 1) if eps * bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        say("sum-saturated", bits, n)
        col = am.Column.from_host_words(np.full(words, word, dtype=np.uint64), n, bits, 1)
        got = int(am.u64(am.sum(col, ws=ws))[0])
        assert got == n * ((1 << bits) - 1), (bits, n, got)
    # bitstreams (SWAR over fields at a stride of w): all-ones streams, long
    # enough that each lane reaches its flush bound (K = 2^w words)
    for bits, n in ((3, 3 * (1 << 24) + 7), (5, (1 << 25) + 3), (11, (1 << 26) + 1)):
        words = (n * bits + 63) // 64
        w = np.full(words, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        tail = n * bits - 64 * (words - 1)
        if tail < 64:
            w[-1] = np.uint64((1 << tail) - 1)
        say("sum-saturated bitstream", bits, n)
        col = am.Column.from_host_words(w, n, bits, 2)
        got = int(am.u64(am.sum(col, ws=ws))[0])
        assert got == n * ((1 << bits) - 1), (bits, n, got)


def _width(v):
    return max(1, int(v).bit_length())


def test_groupby(am, oracle_mod):
    rng = np.random.default_rng(9)
    ws = am.Workspace(0)
    for kbits, vbits in ((10, 20), (12, 32), (1, 5), (4, 1)):
        for n in (1, 4097, 100003, 2000003):
            g = 1 << kbits
            # Zipf-like skew: rank r with weight 1/(r+1)^1.1
            w = 1.0 / np.arange(1, g + 1) ** 1.1
            keys = rng.choice(g, size=n, p=w / w.sum()).astype(np.uint64)
            vals = rand_vals(rng, n, vbits)
            for kf, vf in ((1, 1), (2, 2), (4, 3)):
                kc = upload(am, oracle_mod, keys, kbits, kf)
                vc = upload(am, oracle_mod, vals, vbits, vf)
                for mode in (0, 1):
                    for pred in (None, (2, (1 << (vbits - 1)) - 1)):
                        say("groupby", kbits, vbits, n, kf, vf, mode, pred)
                        ok, osum, ocnt, res = am.groupby(kc, vc, mode=mode, value_pred=pred, ws=ws)
                        r = am.u64(res)
                        sums, counts, rows = oracle_mod.groupby(kf, kbits, oracle_mod.pack(kf, kbits, keys), vf, vbits,
                                                                oracle_mod.pack(vf, vbits, vals), n, pred)
                        present = np.nonzero(counts)[0]
                        ng = len(present)
                        exp_s = sums[present]
                        if mode == 1:
                            exp_s = exp_s & np.uint64((1 << vbits) - 1)
                        sb = vbits if mode == 1 else _width(exp_s.max() if ng else 0)
                        cb = _width(counts.max())
                        assert list(map(int, r)) == [ng, sb, cb, rows], (list(map(int, r)), ng, sb, cb, rows)
                        assert np.array_equal(oracle_mod.unpack(1, kbits, am.u64(ok), ng), present.astype(np.uint64))
                        assert np.array_equal(oracle_mod.unpack(1, sb, am.u64(osum), ng), exp_s)
                        assert np.array_equal(oracle_mod.unpack(1, cb, am.u64(ocnt), ng), counts[present])
                        # dead bits of the last packed word are zero
                        for buf, bits in ((ok, kbits), (osum, sb), (ocnt, cb)):
                            wds = oracle_mod.packed_words(1, bits, ng)
                            assert np.array_equal(am.u64(buf)[:wds], oracle_mod.pack(1, bits, oracle_mod.unpack(
                                1, bits, am.u64(buf), ng)))


def _pack_matrix(oracle_mod, X, bits):
    return np.concatenate([oracle_mod.pack(1, bits, X[r]) for r in range(X.shape[0])])


def test_groupby_folds(am, oracle_mod):
    """Enough rows that every block of the group-by partials folds its packed
    LDS table (12-bit keys, 32-bit values: a fold every 32768 rows)."""
    rng = np.random.default_rng(19)
    ws = am.Workspace(0)
    kbits, vbits, n = 12, 32, 9000001
    keys = rng.integers(0, 1 << kbits, size=n, dtype=np.uint64)
    vals = rand_vals(rng, n, vbits)
    kc = upload(am, oracle_mod, keys, kbits, 1)
    vc = upload(am, oracle_mod, vals, vbits, 1)
    say("groupby-folds", kbits, vbits, n)
    ok, osum, ocnt, res = am.groupby(kc, vc, mode=0, ws=ws)
    r = am.u64(res)
    sums = np.zeros(1 << kbits, np.uint64)
    np.add.at(sums, keys.astype(np.int64), vals)
    counts = np.bincount(keys.astype(np.int64), minlength=1 << kbits).astype(np.uint64)
    present = np.nonzero(counts)[0]
    ng = len(present)
    assert int(r[0]) == ng and int(r[3]) == n
    assert np.array_equal(oracle_mod.unpack(1, kbits, am.u64(ok), ng), present.astype(np.uint64))
    assert np.array_equal(oracle_mod.unpack(1, int(r[1]), am.u64(osum), ng), sums[present])
    assert np.array_equal(oracle_mod.unpack(1, int(r[2]), am.u64(ocnt), ng), counts[present])


def test_matmul_packed(am, oracle_mod):
    import torch
    rng = np.random.default_rng(10)
    shapes = [(1, 1, 1), (16, 16, 64), (17, 33, 65), (128, 128, 64), (130, 200, 300), (257, 129, 1000)]
    for bits in (7, 6, 5, 4, 3, 2, 1):  # 63, 60 and 64 payload bits per word
        probs, exps = [], []
        for (M, N, K) in shapes:
            A = rng.integers(0, 1 << bits, (M, K)).astype(np.uint64)
            B = rng.integers(0, 1 << bits, (K, N)).astype(np.uint64)
            pa, pb = _pack_matrix(oracle_mod, A, bits), _pack_matrix(oracle_mod, B, bits)
            exps.append((oracle_mod.matmul_packed(bits, pa, pb, M, N, K), A, B))
            probs.append((torch.from_numpy(pa.view(np.int64)).cuda(), torch.from_numpy(pb.view(np.int64)).cuda(), M, N, K))
        say("matmul batch", bits, shapes)
        outs = am.matmul_packed_batched(bits, probs, want_i32=True)
        torch.cuda.synchronize()
        for (cp, ci, ob), (C, A, B), (_, _, M, N, K) in zip(outs, exps, probs):
            assert np.array_equal(C, A.astype(np.int64) @ B.astype(np.int64))
            got = ci.cpu().numpy().reshape(M, N).astype(np.int64)
            assert np.array_equal(got, C), (bits, M, N, K)
            assert ob == _width(K * ((1 << bits) - 1) ** 2)
            exp_words = _pack_matrix(oracle_mod, C.astype(np.uint64), ob)
            assert np.array_equal(am.u64(cp)[:len(exp_words)], exp_words), (bits, M, N, K, ob)


def test_matmul_single_large(am, oracle_mod):
    """One matmul with M, N, K >= 512 (the MFMA GEMM with the fused epilogue
    packer or the separate row packer), with and without the int32 copy;
    ragged N and K."""
    import torch
    rng = np.random.default_rng(13)
    for bits, (M, N, K), want in ((7, (520, 777, 600), True), (5, (1024, 513, 2000), False),
                                  (1, (512, 640, 4100), True)):
        A = rng.integers(0, 1 << bits, (M, K)).astype(np.uint64)
        B = rng.integers(0, 1 << bits, (K, N)).astype(np.uint64)
        pa, pb = _pack_matrix(oracle_mod, A, bits), _pack_matrix(oracle_mod, B, bits)
        # entries < 2^31 are exact in float64 products summed in float64
        C = np.rint(A.astype(np.float64) @ B.astype(np.float64)).astype(np.int64)
        if M * N * K <= 520 * 777 * 600:
            assert np.array_equal(oracle_mod.matmul_packed(bits, pa, pb, M, N, K), C)
        say("matmul single", bits, M, N, K, want)
        cp, ci, ob = am.matmul_packed_batched(
            bits, [(torch.from_numpy(pa.view(np.int64)).cuda(), torch.from_numpy(pb.view(np.int64)).cuda(), M, N, K)],
            want_i32=want)[0]
        torch.cuda.synchronize()
        if want:
            assert np.array_equal(ci.cpu().numpy().reshape(M, N).astype(np.int64), C), (bits, M, N, K)
        assert ob == _width(K * ((1 << bits) - 1) ** 2)
        exp_words = _pack_matrix(oracle_mod, C.astype(np.uint64), ob)
        assert np.array_equal(am.u64(cp)[:len(exp_words)], exp_words), (bits, M, N, K, ob)


def test_matmul_uneven_batch_packed_only(am, oracle_mod):
    """A batch of uneven shapes with packed output only (row packer path)."""
    import torch
    rng = np.random.default_rng(11)
    probs, exps = [], []
    for _ in range(12):
        M, N, K = (int(x) for x in rng.integers(64, 400, 3))
        A = rng.integers(0, 128, (M, K)).astype(np.uint64)
        B = rng.integers(0, 128, (K, N)).astype(np.uint64)
        pa, pb = _pack_matrix(oracle_mod, A, 7), _pack_matrix(oracle_mod, B, 7)
        exps.append(A.astype(np.int64) @ B.astype(np.int64))
        probs.append((torch.from_numpy(pa.view(np.int64)).cuda(), torch.from_numpy(pb.view(np.int64)).cuda(), M, N, K))
    say("matmul uneven batch")
    outs = am.matmul_packed_batched(7, probs)
    torch.cuda.synchronize()
    for (cp, _, ob), C in zip(outs, exps):
        exp_words = _pack_matrix(oracle_mod, C.astype(np.uint64), ob)
        assert np.array_equal(am.u64(cp)[:len(exp_words)], exp_words)


def test_pipeline_batch(am, oracle_mod):
    """Fused filter -> compressstore(a+b) -> sum over a batch of uneven columns."""
    rng = np.random.default_rng(12)
    cols, refs = [], []
    for i in range(24):
        n = int([1, 63, 4096, 4097, 30011, 200003][i % 6] * (1 + (i // 6) % 2))
        w = int(rng.integers(3, 32))
        kind = i % 3
        if kind == 0:
            a = rand_vals(rng, n, w)
        elif kind == 1:
            g = min(1 << w, 4096)
            p = 1.0 / np.arange(1, g + 1) ** 1.2
            a = rng.choice(g, size=n, p=p / p.sum()).astype(np.uint64)
        else:
            a = np.minimum((np.arange(n) * ((1 << w) - 1) // max(n, 1)) + rng.integers(0, 8, n), (1 << w) - 1)
            a = a.astype(np.uint64)
        b = rand_vals(rng, n, w)
        c = int(rng.integers(0, 1 << w))
        fmt = (1, 2)[i % 2]
        sb = w + 1 if i % 4 else max(1, w - 1)  # some columns too narrow: overflow must be reported
        cols.append((upload(am, oracle_mod, a, w, 1), upload(am, oracle_mod, b, w, 1), (0, c), fmt, sb))
        v = (a + b) & np.uint64((1 << (w + 1)) - 1)
        sel = a < np.uint64(c)
        refs.append((v, sel, fmt, sb))
    say("pipeline batch", len(cols))
    batch = am.PipelineBatch(cols)
    for rep in range(3):  # repeated runs on one workspace (epoch tags)
        batch.run()
    import torch
    torch.cuda.synchronize()
    outs, st = batch.results()
    for i, ((v, sel, fmt, sb), out) in enumerate(zip(refs, outs)):
        n = len(v)
        mask = np.packbits(sel, bitorder="little")
        mask = np.concatenate([mask, np.zeros((-len(mask)) % 8, np.uint8)]).view(np.uint64)
        k, exp_words, ovf = oracle_mod.compressstore(0, 64, v, n, mask, fmt, sb)
        total = int(v[sel].sum(dtype=np.uint64))
        assert [int(x) for x in st[i][:3]] == [k, ovf, total], (i, list(st[i]), k, ovf, total)
        used = oracle_mod.packed_words(fmt, sb, k)
        assert np.array_equal(out.host_words()[:used], exp_words[:used]), i


def test_planner_kernels(am, oracle_mod, tmp_path):
    """k_plan_stats features are sane and k_plan_mlp equals a numpy forward pass."""
    import torch
    from amethyst_amd import planner
    cols = []
    for seed in range(6):
        spec = planner.column_spec(seed, nmax=1 << 18)
        cols.append(planner.make_inputs(seed, spec))
    f = planner.features(cols)
    fr = f.cpu().numpy()
    for (a, b, pred), row in zip(cols, fr):
        assert abs(row[0] - a.bits / 32.0) < 1e-6
        assert 0.0 <= row[2] <= 1.0 and abs(row[3:19].sum() - 1.0) < 1e-4
    rng = np.random.default_rng(0)
    P = planner.NPLANS
    nw = 64 * 32 + 64 + 64 * 64 + 64 + P * 64 + P
    z = {"weights": (rng.standard_normal(nw) * 0.2).astype(np.float32), "mean": fr.mean(0).astype(np.float32),
         "std": (fr.std(0) + 1e-3).astype(np.float32), "nplans": np.int32(P)}
    path = tmp_path / "w.npz"
    np.savez(path, **z)
    pl = planner.Planner(str(path), min_margin=0.0)
    choice, sc = pl.choose(f, scores=True)
    ref = planner.mlp_reference(np.load(path), fr)
    assert np.allclose(sc, ref, rtol=1e-4, atol=1e-4)
    assert np.array_equal(choice, ref.argmin(axis=1))
    # an impossible margin sends every column to the analytic plan
    pl.min_margin = 1e9
    choice2, _ = pl.choose(f)
    assert (choice2 == planner.ANALYTIC).all()


def test_plain_chain_matches_fused(am, oracle_mod):
    """The reference-style plain chain and the fused batch agree (totals)."""
    import torch
    from amethyst_amd import planner
    cols = [planner.make_inputs(s, planner.column_spec(s, nmax=1 << 17)) for s in range(7)]
    plain = planner.run_plain(cols)
    f = planner.features(cols).cpu().numpy()
    for p in range(planner.NPLANS):
        outs, stats, _, _ = planner.run_batch(cols, [p] * len(cols), f)
        assert [int(s[2]) for s in stats] == plain, p


def test_gen_masks_match_oracle(am, oracle_mod):
    """Bernoulli and clustered (two-state chain) masks, bit-exact against the
    oracle's serial restatement, ragged lengths around the 64-bit word and
    the 4096-bit chain chunk."""
    for n in (1, 63, 64, 4095, 4096, 4097, 100003, 1 << 20):
        for s in (0.0, 0.01, 0.5, 0.99, 1.0):
            say("gen_mask_bernoulli", n, s)
            got = am.u64(am.gen_mask_bernoulli(11, n, s))[:(n + 63) // 64]
            assert np.array_equal(got, oracle_mod.gen_mask_bernoulli(11, n, s)), (n, s)
        for s in (0.01, 0.5, 0.99):
            for mean in (256.0, 3.0):
                say("gen_mask_runs", n, s, mean)
                got = am.u64(am.gen_mask_runs(12, n, s, mean))[:(n + 63) // 64]
                assert np.array_equal(got, oracle_mod.gen_mask_runs(12, n, s, mean)), (n, s, mean)


def test_gen_reference_masks(am, oracle_mod):
    """The reference driver's generate_mask_uniform and create_bitmask bytes,
    produced on the GPU, equal the oracle's restatements (pinned to golden
    bytes of the reference's generator)."""
    for nbytes in (1, 7, 8, 4099, 1 << 18):
        for s in (0.01, 0.5, 0.9):
            say("gen_mask_uniform_ref", nbytes, s)
            got = am.u64(am.gen_mask_uniform_ref(nbytes, s)).view(np.uint8)[:nbytes]
            assert np.array_equal(got, oracle_mod.mask_uniform_ref(nbytes, s)), (nbytes, s)
    for total in (8, 64, 1000, 1 << 20, (1 << 20) + 24):
        for s, k in ((0.01, 1), (0.5, 8), (0.5, 1), (0.99, 8), (0.3, 3)):
            say("gen_mask_clusters_ref", total, s, k)
            got = am.u64(am.gen_mask_clusters_ref(s, k, total)).view(np.uint8)[:total // 8]
            assert np.array_equal(got, oracle_mod.mask_clusters_ref(s, k, total)), (total, s, k)


def _zipf_thresholds(g, s):
    p = 1.0 / np.arange(1, g + 1) ** s
    c = np.cumsum(p / p.sum())
    thr = np.minimum(np.floor(c * 4294967296.0), 4294967295).astype(np.uint64)
    thr[-1] = 4294967295
    return thr.astype(np.uint32)


def test_gen_from_cdf_and_sorted_noise(am, oracle_mod):
    """Inverse-CDF draws (the Zipf keys of the group-by and uneven-batch
    workloads) and sorted-with-noise columns, every format, against the
    oracle. One threshold table holds exact draw values, so a `<=` for `<`
    in the search would change those elements."""
    n = 40009
    draws = oracle_mod.prng(5, 64)
    exact = np.unique(np.concatenate([draws[:20], [4294967295]]).astype(np.uint64)).astype(np.uint32)
    tables = ((_zipf_thresholds(1024, 1.1), 10), (_zipf_thresholds(4096, 1.2), 12), (exact, 5),
              (np.array([4294967295], np.uint32), 3))
    for thr, bits in tables:
        exp = oracle_mod.gen_from_cdf(5, thr, bits, n)
        for fmt in FORMATS:
            say("gen_from_cdf", len(thr), bits, fmt)
            col = am.gen_from_cdf(5, thr, n, bits, fmt)
            assert np.array_equal(col.host_words(), oracle_mod.pack(fmt, bits, exp)), (len(thr), fmt)
    for bits, nb in ((3, 1), (11, 5), (20, 14), (31, 25), (40, 32)):
        exp = oracle_mod.gen_sorted_noise(9, bits, nb, n)
        for fmt in FORMATS:
            if not fmt_ok(fmt, bits):
                continue
            say("gen_sorted_noise", bits, nb, fmt)
            col = am.gen_sorted_noise(9, n, bits, fmt, nb)
            assert np.array_equal(col.host_words(), oracle_mod.pack(fmt, bits, exp)), (bits, fmt)


def test_compress_clustered_masks(am, oracle_mod):
    """Config C2's clustered masks (alternating runs, 1-runs geometric with
    mean 256) and the reference driver's own mask generators (uniform bytes
    and 1 or 8 equal clusters, create_bitmask), driving compaction into slab
    (same and narrower width) and u32 outputs, against the oracle."""
    n = 3_000_017
    ws = am.Workspace(n)
    masks = []
    for s in (0.01, 0.5, 0.99):
        masks.append((f"runs{s}", am.gen_mask_runs(21, n, s, 256.0)))
    masks.append(("ref-uniform0.5", am.gen_mask_uniform_ref(((n + 63) // 64) * 8, 0.5)))
    masks.append(("ref-1cluster0.01", am.gen_mask_clusters_ref(0.01, 1, n)))
    masks.append(("ref-8clusters0.5", am.gen_mask_clusters_ref(0.5, 8, n)))
    for w in (4, 8, 11, 17, 24, 32):
        vals = oracle_mod.prng(100 + w, n).astype(np.uint64) & np.uint64((1 << w) - 1)
        for name, m in masks:
            mw = am.u64(m)[:(n + 63) // 64].copy()
            for fo, ob in ((am.SLAB, w), (am.U32, w), (am.SLAB, max(1, w - 3))):
                say("compress-clustered", w, name, fo, ob)
                _check_compress(am, oracle_mod, vals, w, 1, ob, fo, mask_words=mw, ws=ws)
    ws.check()


def test_config_c3_exact(am, oracle_mod):
    """Config C3 as specified: 2^28 rows, keys 10-bit Zipf(1.1) over 1024
    ranks (inverse CDF of the PCG stream, seed 42), values 20-bit uniform
    (seed 43), both slab-packed; exact u64 SUM + COUNT, key-sorted,
    minimal-width output; and the chained filter v > 2^19 -> group-by.
    Expected sums and counts from np.bincount over the host-decoded draws."""
    n = 1 << 28
    thr = _zipf_thresholds(1024, 1.1)
    keys = am.gen_from_cdf(42, thr, n, 10, am.SLAB)
    vals = am.gen_uniform(43, n, 20, am.SLAB)
    k = _cdf_rank(thr, oracle_mod.prng(42, n))
    v = (oracle_mod.prng(43, n) & np.uint32(0xFFFFF)).astype(np.float64)
    ws = am.Workspace(0)
    for pred in (None, (am.GT, 1 << 19)):
        say("config-c3", pred)
        sel = slice(None) if pred is None else (v > float(1 << 19))
        kk = k if pred is None else k[sel]
        vv = v if pred is None else v[sel]
        cnt = np.bincount(kk, minlength=1024).astype(np.uint64)
        sm = np.bincount(kk, weights=vv, minlength=1024)  # < 2^53: exact in float64
        sm = sm.astype(np.uint64)
        ok, osum, ocnt, res = am.groupby(keys, vals, value_pred=pred, ws=ws)
        r = am.u64(res)
        present = np.nonzero(cnt)[0]
        ng = len(present)
        assert int(r[0]) == ng and int(r[3]) == int(cnt.sum())
        assert int(r[1]) == _width(int(sm.max())) and int(r[2]) == _width(int(cnt.max()))
        assert np.array_equal(oracle_mod.unpack(1, 10, am.u64(ok), ng), present.astype(np.uint64))
        assert np.array_equal(oracle_mod.unpack(1, int(r[1]), am.u64(osum), ng), sm[present])
        assert np.array_equal(oracle_mod.unpack(1, int(r[2]), am.u64(ocnt), ng), cnt[present])
        del ok, osum, ocnt, res
    ws.check()


def _cdf_rank(thr, d):
    """First b with d < thr[b] (capped at the last bin), i.e. searchsorted
    'right', through a table over the top 16 bits of d; draws whose
    65536-wide bucket holds a threshold take the exact search."""
    top = np.arange(1 << 16, dtype=np.uint64) << np.uint64(16)
    lo = np.searchsorted(thr, top, side="right").astype(np.int16)
    hi = np.searchsorted(thr, top + np.uint64(65535), side="right").astype(np.int16)
    amb_bucket = hi != lo
    t = d >> np.uint32(16)
    k = lo[t]
    amb = np.nonzero(amb_bucket[t])[0]
    k[amb] = np.searchsorted(thr, d[amb], side="right")
    return np.minimum(k, len(thr) - 1)


def _pack_rows_u7(X):
    """Row-padded 7-bit slab matrix (9 fields per word) by numpy."""
    R, C = X.shape
    cp = -(-C // 9) * 9
    Y = np.zeros((R, cp), np.uint64)
    Y[:, :C] = X
    sh = (np.arange(9, dtype=np.uint64) * np.uint64(7))
    return (Y.reshape(R, cp // 9, 9) << sh).sum(axis=2, dtype=np.uint64).reshape(-1)


def _unpack_rows(words, R, C, bits):
    eps = 64 // bits
    wpr = -(-C // eps)
    W = np.asarray(words[:R * wpr], np.uint64).reshape(R, wpr, 1)
    sh = np.arange(eps, dtype=np.uint64) * np.uint64(bits)
    return ((W >> sh) & np.uint64((1 << bits) - 1)).reshape(R, wpr * eps)[:, :C]


def test_config_c5_full_size_sampled(am, oracle_mod):
    """Config C5 at full size: one 8192^3 product of 7-bit operands and a
    batch of 256 products with M, N, K ~ U[64, 2048]; 4096 sampled entries
    of the big product and 16 per batch member are checked against int64
    dot products on the host, in both the int32 and the packed output."""
    import torch
    rng = np.random.default_rng(55)
    M = N = K = 8192
    A = rng.integers(0, 128, (M, K), dtype=np.uint64)
    B = rng.integers(0, 128, (K, N), dtype=np.uint64)
    assert np.array_equal(_pack_rows_u7(A[:3, :100]), _pack_matrix(oracle_mod, A[:3, :100], 7))
    pa, pb = _pack_rows_u7(A), _pack_rows_u7(B)
    say("c5 8192^3")
    cp, ci, ob = am.matmul_packed_batched(
        7, [(torch.from_numpy(pa.view(np.int64)).cuda(), torch.from_numpy(pb.view(np.int64)).cuda(), M, N, K)],
        want_i32=True)[0]
    torch.cuda.synchronize()
    assert ob == 27
    i = rng.integers(0, M, 4096)
    j = rng.integers(0, N, 4096)
    exp = np.einsum("sk,sk->s", A[i].astype(np.int64), B[:, j].T.astype(np.int64))
    C32 = ci.view(M, N)
    got = C32[torch.from_numpy(i).cuda(), torch.from_numpy(j).cuda()].cpu().numpy().astype(np.int64)
    assert np.array_equal(got, exp)
    wpr = -(-N // (64 // ob))
    rows = np.unique(i[:64])
    P = _unpack_rows(am.u64(cp.view(-1)[:M * wpr].view(M, wpr)[torch.from_numpy(rows).cuda()].reshape(-1)),
                     len(rows), N, ob)
    assert np.array_equal(P.astype(np.int64), C32[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.int64))
    del cp, ci, C32
    probs, host = [], []
    for t in range(256):
        m, nn, k = (int(x) for x in rng.integers(64, 2049, 3))
        a = rng.integers(0, 128, (m, k), dtype=np.uint64)
        b = rng.integers(0, 128, (k, nn), dtype=np.uint64)
        probs.append((torch.from_numpy(_pack_rows_u7(a).view(np.int64)).cuda(),
                      torch.from_numpy(_pack_rows_u7(b).view(np.int64)).cuda(), m, nn, k))
        host.append((a, b))
    say("c5 batch of 256")
    outs = am.matmul_packed_batched(7, probs, want_i32=True)
    torch.cuda.synchronize()
    for (cp, ci, ob), (a, b), (_, _, m, nn, k) in zip(outs, host, probs):
        ii = rng.integers(0, m, 16)
        jj = rng.integers(0, nn, 16)
        exp = np.einsum("sk,sk->s", a[ii].astype(np.int64), b[:, jj].T.astype(np.int64))
        Cm = ci.view(m, nn).cpu().numpy().astype(np.int64)
        assert np.array_equal(Cm[ii, jj], exp), (m, nn, k)
        assert np.array_equal(_unpack_rows(am.u64(cp), m, nn, ob).astype(np.int64), Cm), (m, nn, k)


def _pipeline_cols(am, oracle_mod, rng, ncols, fmt_in):
    cols, refs = [], []
    for i in range(ncols):
        n = int([1, 63, 2048, 4096, 4097, 30011, 200003][i % 7] * (1 + (i // 7) % 2))
        w = int(rng.integers(3, 32))
        kind = i % 3
        if kind == 0:
            a = rand_vals(rng, n, w)
        elif kind == 1:
            g = min(1 << w, 4096)
            p = 1.0 / np.arange(1, g + 1) ** 1.2
            a = rng.choice(g, size=n, p=p / p.sum()).astype(np.uint64)
        else:
            a = np.minimum((np.arange(n) * ((1 << w) - 1) // max(n, 1)) + rng.integers(0, 8, n), (1 << w) - 1)
            a = a.astype(np.uint64)
        b = rand_vals(rng, n, w)
        c = int(rng.integers(0, 1 << w))
        fmt = (1, 2)[i % 2]
        sb = w + 1 if i % 4 else max(1, w - 1)  # some columns too narrow: overflow must be reported
        cols.append((upload(am, oracle_mod, a, w, fmt_in), upload(am, oracle_mod, b, w, fmt_in), (0, c), fmt, sb))
        v = (a + b) & np.uint64((1 << (w + 1)) - 1)
        refs.append((v, a < np.uint64(c), fmt, sb))
    return cols, refs


def _check_pipeline(oracle_mod, refs, outs, st, tag):
    for i, ((v, sel, fmt, sb), out) in enumerate(zip(refs, outs)):
        n = len(v)
        mask = np.packbits(sel, bitorder="little")
        mask = np.concatenate([mask, np.zeros((-len(mask)) % 8, np.uint8)]).view(np.uint64)
        k, exp_words, ovf = oracle_mod.compressstore(0, 64, v, n, mask, fmt, sb)
        total = int(v[sel].sum(dtype=np.uint64))
        assert [int(x) for x in st[i][:3]] == [k, ovf, total], (tag, i, list(st[i]), k, ovf, total)
        used = oracle_mod.packed_words(fmt, sb, k)
        assert np.array_equal(out.host_words()[:used], exp_words[:used]), (tag, i)


@pytest.mark.parametrize("fmt_in", [1, 2, 4, 3])
@pytest.mark.parametrize("variant", [8, 16, 32])
def test_pipeline_variants_and_formats(am, oracle_mod, fmt_in, variant):
    """The fused chain for every input decoder (slab, bitstream, u32, tight)
    and every tile size: packed s words, counts, overflow and totals."""
    import torch
    rng = np.random.default_rng(100 + 10 * fmt_in + variant)
    cols, refs = _pipeline_cols(am, oracle_mod, rng, 14, fmt_in)
    say("pipeline", fmt_in, variant)
    batch = am.PipelineBatch(cols, variant=variant)
    batch.run()
    batch.run()
    torch.cuda.synchronize()
    outs, st = batch.results()
    _check_pipeline(oracle_mod, refs, outs, st, (fmt_in, variant))


def test_pipeline_slab_every_width_and_predicate(am, oracle_mod):
    """Slab inputs of every width 1..31 (2..31 take the word-per-lane
    decode, 1 the per-element one) with all six predicates, every tile size,
    lengths that end mid-word and mid-tile, values at the edges of the
    range, in one batch: packed s words, counts, overflow and totals."""
    import torch
    rng = np.random.default_rng(2024)
    ops = {0: np.less, 1: np.less_equal, 2: np.greater, 3: np.greater_equal, 4: np.equal, 5: np.not_equal}
    cols, refs, variants = [], [], []
    for w in range(1, 32):
        n = int(rng.integers(1, 70000)) if w % 5 else int(rng.integers(8192 * 3, 8192 * 5))
        a = rand_vals(rng, n, w)
        a[rng.random(n) < 0.05] = (1 << w) - 1  # the top value and zero are common predicate edges
        a[rng.random(n) < 0.05] = 0
        b = rand_vals(rng, n, w)
        op = int(rng.integers(0, 6))
        c = int(rng.integers(0, 1 << w)) if op < 4 else int(a[rng.integers(0, n)])
        fmt = (1, 2)[w % 2]
        sb = w + 1 if w % 3 else max(1, w - 1)
        cols.append((upload(am, oracle_mod, a, w, 1), upload(am, oracle_mod, b, w, 1), (op, c), fmt, sb))
        v = (a + b) & np.uint64((1 << (w + 1)) - 1)
        refs.append((v, ops[op](a, np.uint64(c)), fmt, sb))
        variants.append((8, 16, 32)[w % 3])
    say("pipeline slab widths 1..31")
    batch = am.PipelineBatch(cols, variant=variants)
    batch.run()
    batch.run()
    torch.cuda.synchronize()
    outs, st = batch.results()
    _check_pipeline(oracle_mod, refs, outs, st, "slab widths")


def _pipeline_xcd_local(batch):
    """The batch record's layout flag: word 3 of the workspace holds the
    run's block count (low half) and whether each column runs on one XCD."""
    return int(batch.ws[3].item()) >> 32


def test_pipeline_xcd_local_and_fallback_layouts(am, oracle_mod):
    """Many columns of similar length take the XCD-local block map (columns
    whole on one XCD, padding blocks idle); a batch whose one column holds
    most tiles keeps the column-by-column map. Both against the oracle,
    run twice each on one workspace."""
    import torch
    rng = np.random.default_rng(515)
    ops = {0: np.less, 2: np.greater}
    for ncols, big in ((48, False), (6, True)):
        cols, refs, variants = [], [], []
        for i in range(ncols):
            n = int(rng.integers(30000, 60000)) if not (big and i == 0) else 900001
            w = int(rng.integers(2, 32))
            a = rand_vals(rng, n, w)
            b = rand_vals(rng, n, w)
            op = (0, 2)[i % 2]
            c = int(rng.integers(0, 1 << w))
            fin = (1, 1, 2, 4)[i % 4]
            fmt = (1, 2)[i % 2]
            sb = w + 1 if i % 5 else max(1, w - 1)
            cols.append((upload(am, oracle_mod, a, w, fin), upload(am, oracle_mod, b, w, fin), (op, c), fmt, sb))
            v = (a + b) & np.uint64((1 << (w + 1)) - 1)
            refs.append((v, ops[op](a, np.uint64(c)), fmt, sb))
            variants.append((8, 16, 32)[i % 3])
        say("pipeline layout", ncols, big)
        batch = am.PipelineBatch(cols, variant=variants)
        assert _pipeline_xcd_local(batch) == (0 if big else 1), (ncols, big)
        for _ in range(2):
            batch.stats.zero_()
            batch.run()
            torch.cuda.synchronize()
            outs, st = batch.results()
            _check_pipeline(oracle_mod, refs, outs, st, ("layout", ncols, big))


def test_pipeline_graph_replay(am, oracle_mod):
    """Config C4's launch form: the batch captured once in a graph and
    replayed; before every replay the outputs and statistics are cleared
    eagerly (and other eager work runs), and every replay is checked against
    the oracle."""
    import torch
    rng = np.random.default_rng(77)
    cols, refs = _pipeline_cols(am, oracle_mod, rng, 21, 1)
    batch = am.PipelineBatch(cols, variant=16)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batch.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.run()
    for rep in range(4):
        for o in batch.outs:
            o.buf.zero_()
        batch.stats.zero_()
        junk = am.gen_uniform(rep, 100000, 11, am.SLAB)  # eager work between replays
        torch.cuda.synchronize()
        say("pipeline graph replay", rep)
        g.replay()
        torch.cuda.synchronize()
        outs, st = batch.results()
        _check_pipeline(oracle_mod, refs, outs, st, ("replay", rep))
        del junk


def _check_groupby(am, oracle_mod, keys, vals, kbits, vbits, kf, vf, mode, pred, ws, tag):
    n = len(keys)
    kc = upload(am, oracle_mod, keys, kbits, kf)
    vc = upload(am, oracle_mod, vals, vbits, vf)
    ok, osum, ocnt, res = am.groupby(kc, vc, mode=mode, value_pred=pred, ws=ws)
    r = am.u64(res)
    sums, counts, rows = oracle_mod.groupby(kf, kbits, oracle_mod.pack(kf, kbits, keys), vf, vbits,
                                            oracle_mod.pack(vf, vbits, vals), n, pred)
    present = np.nonzero(counts)[0]
    ng = len(present)
    exp_s = sums[present]
    if mode == 1:
        exp_s = exp_s & np.uint64((1 << vbits) - 1)
    sb = vbits if mode == 1 else _width(exp_s.max() if ng else 0)
    cb = _width(counts.max())
    assert list(map(int, r)) == [ng, sb, cb, rows], (tag, list(map(int, r)), ng, sb, cb, rows)
    assert np.array_equal(oracle_mod.unpack(1, kbits, am.u64(ok), ng), present.astype(np.uint64)), tag
    assert np.array_equal(oracle_mod.unpack(1, sb, am.u64(osum), ng), exp_s), tag
    assert np.array_equal(oracle_mod.unpack(1, cb, am.u64(ocnt), ng), counts[present]), tag
    for buf, bits in ((ok, kbits), (osum, sb), (ocnt, cb)):  # dead and padding bits zero
        wds = oracle_mod.packed_words(1, bits, ng)
        assert np.array_equal(am.u64(buf)[:wds], oracle_mod.pack(1, bits, oracle_mod.unpack(1, bits, am.u64(buf), ng)))


def test_groupby_wide_keys(am, oracle_mod):
    """Keys of 13..24 bits (partitioned path) against the oracle: uniform keys
    (partitions of one aggregation chunk, stored directly) and Zipf(1.1)
    ranks (a few partitions of many chunks, merged with atomics), slab /
    bitstream / container inputs, both modes, with and without the chained
    value predicate, plus an empty selection."""
    rng = np.random.default_rng(31)
    for kbits, vbits, n, dist in ((13, 20, 100003, "zipf"), (16, 20, 2000003, "zipf"), (16, 32, 300007, "uniform"),
                                  (20, 1, 50000, "uniform"), (24, 20, 3000017, "uniform"), (24, 32, 1000003, "zipf"),
                                  (18, 11, 1, "uniform")):
        g = 1 << kbits
        if dist == "zipf":
            p = 1.0 / np.arange(1, g + 1) ** 1.1
            keys = rng.choice(g, size=n, p=p / p.sum()).astype(np.uint64)
        else:
            keys = rng.integers(0, g, size=n, dtype=np.uint64)
        vals = rand_vals(rng, n, vbits)
        ws = am.Workspace.for_groupby(n, kbits, vbits)
        for kf, vf in ((1, 1), (2, 2), (4, 3)):
            for mode in (0, 1):
                for pred in (None, (2, (1 << (vbits - 1)) - 1), (0, 0)):
                    if pred is not None and (mode, kf) != (0, 1):
                        continue
                    say("groupby-wide", kbits, vbits, n, dist, kf, vf, mode, pred)
                    _check_groupby(am, oracle_mod, keys, vals, kbits, vbits, kf, vf, mode, pred, ws,
                                   (kbits, vbits, n, kf, vf, mode, pred))


def test_pipeline_mixed_tiles_one_launch(am, oracle_mod):
    """Columns of every tile size and several input formats' worth of widths
    in ONE launch, run repeatedly on one workspace; then a second batch
    prepared on the same workspace."""
    import torch
    rng = np.random.default_rng(91)
    cols, refs = _pipeline_cols(am, oracle_mod, rng, 18, 2)
    variants = [(8, 16, 32, 0)[i % 4] for i in range(len(cols))]
    batch = am.PipelineBatch(cols, variant=variants)
    for _ in range(3):
        batch.run()
    torch.cuda.synchronize()
    outs, st = batch.results()
    _check_pipeline(oracle_mod, refs, outs, st, "mixed")
    cols2, refs2 = _pipeline_cols(am, oracle_mod, rng, 9, 1)
    b2 = am.PipelineBatch(cols2, variant=8)
    b2.ws, b2.ws_bytes = batch.ws, batch.ws_bytes  # reuse the first batch's workspace
    import ctypes
    tt = ctypes.c_uint32(0)
    assert am.lib().am_pipeline_prepare(b2.arr, b2.n, ctypes.c_void_p(b2.ws.data_ptr()), b2.ws_bytes,
                                        ctypes.byref(tt), am._stream()) == 0
    b2.total_tiles = tt.value
    b2.run()
    b2.run()
    torch.cuda.synchronize()
    outs2, st2 = b2.results()
    _check_pipeline(oracle_mod, refs2, outs2, st2, "second batch")


def test_pipeline_run_rejects_mismatched_launch(am, oracle_mod):
    """A run whose tile count or workspace size differs from the prepared
    batch writes stats[0] = UINT64_MAX for every column and nothing else
    (the kernel checks the recorded batch; every block exits normally)."""
    import torch
    rng = np.random.default_rng(92)
    cols, refs = _pipeline_cols(am, oracle_mod, rng, 5, 1)
    batch = am.PipelineBatch(cols, variant=16)
    map_end = 256 + 65536 + (8 * batch.total_tiles + 255) // 256 * 256  # header, small region, block map
    assert am.lib().am_pipeline_run(am.ctypes.c_void_p(batch.ws.data_ptr()), map_end - 8, batch.total_tiles,
                                    am._stream()) == 4  # AM_E_WORKSPACE: the map itself does not fit
    for tiles, wsb in ((batch.total_tiles + 1, batch.ws_bytes), (max(1, batch.total_tiles - 1), batch.ws_bytes),
                       (batch.total_tiles, map_end)):  # less than the batch's job table and descriptors
        batch.stats.zero_()
        st = am.lib().am_pipeline_run(am.ctypes.c_void_p(batch.ws.data_ptr()), wsb, tiles, am._stream())
        assert st == 0
        torch.cuda.synchronize()
        s = batch.results()[1]
        assert all(int(x[0]) == 2**64 - 1 for x in s), (tiles, wsb)
    batch.stats.zero_()
    batch.run()  # the matching launch still works afterwards
    torch.cuda.synchronize()
    outs, st = batch.results()
    _check_pipeline(oracle_mod, refs, outs, st, "after mismatches")


def test_matmul_plan_and_graph_replay(am, oracle_mod):
    """A prepared matmul batch (header uploaded once) run repeatedly and
    replayed from a captured graph, every run checked against numpy int64
    products; and the one-shot entry point on the same problems."""
    import torch
    rng = np.random.default_rng(61)
    probs, exps = [], []
    for (M, N, K) in ((300, 200, 500), (64, 700, 129), (513, 65, 1000)):
        A = rng.integers(0, 128, (M, K)).astype(np.uint64)
        B = rng.integers(0, 128, (K, N)).astype(np.uint64)
        probs.append((torch.from_numpy(_pack_rows_u7(A).view(np.int64)).cuda(),
                      torch.from_numpy(_pack_rows_u7(B).view(np.int64)).cuda(), M, N, K))
        exps.append(A.astype(np.int64) @ B.astype(np.int64))
    plan = am.MatmulPlan(7, probs, want_i32=True)

    def check(tag):
        for (cp, ci, ob), C, (_, _, M, N, K) in zip(plan.outs, exps, probs):
            assert np.array_equal(ci.cpu().numpy().reshape(M, N).astype(np.int64), C), (tag, M, N, K)
            assert np.array_equal(_unpack_rows(am.u64(cp), M, N, ob).astype(np.int64), C), (tag, M, N, K)

    for rep in range(2):
        for cp, ci, _ in plan.outs:
            cp.zero_()
            ci.zero_()
        say("matmul plan run", rep)
        plan.run()
        torch.cuda.synchronize()
        check(("run", rep))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run()
    for rep in range(3):
        for cp, ci, _ in plan.outs:
            cp.zero_()
            ci.zero_()
        torch.cuda.synchronize()
        say("matmul plan graph replay", rep)
        g.replay()
        torch.cuda.synchronize()
        check(("replay", rep))
    outs = am.matmul_packed_batched(7, probs, want_i32=True)
    torch.cuda.synchronize()
    for (cp, ci, ob), C, (_, _, M, N, K) in zip(outs, exps, probs):
        assert np.array_equal(ci.cpu().numpy().reshape(M, N).astype(np.int64), C)


def test_hipcub_baseline_matches_library(am, oracle_mod):
    """The GPU library baselines bench.py times (hipCUB select/reduce over one
    u64 per element; the reference's cub baselines,
    src/benchs/bench_compressstore_baseline_cub.cu:29-35 and
    src/benchs/bench_hop_baseline_cub.cu:32-38) compute what the library's
    compaction and sum compute on the same data, and what numpy computes."""
    import torch
    from amethyst_amd import baseline as hb
    for n, s in ((1, 1.0), (4097, 0.5), (100_003, 0.3)):
        say("hipcub baseline", n, s)
        col = am.gen_uniform(7, n, 11, am.BASELINE)
        mask = am.gen_mask_bernoulli(8, n, s)
        sel = hb.SelectFlagged(col.buf, mask, n)
        sel.run()
        red = hb.ReduceSum(col.buf, n)
        red.run()
        out, stats, _ = am.compressstore(col, 64, am.BASELINE, mask=mask)
        tot = am.sum(col)
        torch.cuda.synchronize()
        vals = col.host_words()
        bits = np.unpackbits(am.u64(mask)[:(n + 63) // 64].view(np.uint8), bitorder="little")[:n].astype(bool)
        exp = vals[bits]
        k = int(am.u64(stats)[0])
        assert k == len(exp) == int(am.u64(sel.count)[0]), (n, k, len(exp))
        assert np.array_equal(am.u64(sel.out)[:k], exp), n
        assert np.array_equal(out.host_words()[:k], exp), n
        assert int(am.u64(red.out)[0]) == int(am.u64(tot)[0]) == int(vals.sum(dtype=np.uint64)), n
