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
    assert np.array_equal(am.u64(mo)[:(n + 63) // 64], mask_words)
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
    """One matmul with M, N, K >= 512 (the library int8 GEMM + row packer path),
    with and without the int32 copy; ragged N and K."""
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
