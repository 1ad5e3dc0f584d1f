"""Pins the CPU oracle (oracle/oracle.c) -- runs without a GPU.

Each check compares the oracle with something independent of it: the
reference's own code (oracle/_ref, built from src/fast_prng.cuh and
src/bit_iterator.cuh), the committed golden vectors made from that build,
numpy closed forms, or invariants/brute force -- chosen so that a plausible
slip (wrong bit order, off-by-one straddle, dropped mask, transposed operand)
fails one of them.
"""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mask_of(bits):
    return np.uint64((1 << bits) - 1) if bits < 64 else M64


def rand_vals(rng, n, bits):
    hi = rng.integers(0, 2**32, size=n, dtype=np.uint64)
    lo = rng.integers(0, 2**32, size=n, dtype=np.uint64)
    return ((hi << np.uint64(32)) | lo) & mask_of(bits)


def lsb_mask_words(sel):
    m = np.packbits(np.asarray(sel, bool), bitorder="little")
    return np.concatenate([m, np.zeros((-len(m)) % 8, np.uint8)]).view(np.uint64)


# ---------------------------------------------------------------- PRNG

def test_prng_golden(oracle_mod):
    for seed, draws in GOLD["prng"].items():
        assert oracle_mod.prng(int(seed), 64).tolist() == draws
    ps = GOLD["prng_skip"]
    assert oracle_mod.prng(ps["seed"], 16, skip=ps["skip"]).tolist() == ps["draws"]


def test_prng_against_reference_build(oracle_mod):
    import ctypes
    R = oracle_mod.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for seed in (42, 0, 99, 2**63 + 5):
        out = np.empty(4096, np.uint32)
        R.amref_prng_fill(seed, 7, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 4096)
        assert np.array_equal(out, oracle_mod.prng(seed, 4096, skip=7))


def test_pcg32_published_algorithm(oracle_mod):
    """PCG32 XSH-RR (O'Neill 2014) with multiplier 6364136223846793005 and
    increment 1, seeded by splitmix64 (Vigna) -- restated here in Python."""
    def splitmix(seed):
        z = (seed + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)

    st = splitmix(42)
    out = []
    for _ in range(32):
        old = st
        st = (old * 6364136223846793005 + 1) & (2**64 - 1)
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((x >> rot) | (x << ((-rot) & 31))) & 0xFFFFFFFF)
    assert oracle_mod.prng(42, 32).tolist() == out


def test_main_words_layout(oracle_mod):
    """src/main.cu:242-250: buffers alternate draws, word = hi<<32|lo."""
    a, b = oracle_mod.main_words(42, 8)
    d = oracle_mod.prng(42, 32).astype(np.uint64)
    exp_a = (d[0::4] << np.uint64(32)) | d[1::4]
    exp_b = (d[2::4] << np.uint64(32)) | d[3::4]
    assert np.array_equal(a, np.where(exp_a == 0, 1, exp_a))
    assert np.array_equal(b, np.where(exp_b == 0, 1, exp_b))


# ---------------------------------------------------------------- masks

def test_bit_iterator_is_lsb_first(oracle_mod):
    bi = GOLD["bit_iterator"]
    words = np.array(bi["bytes"], np.uint8).view(np.uint64)
    # the oracle's mask reading (compressstore) must agree with bit_iterator
    vals = np.arange(64, dtype=np.uint64)
    k, out, _ = oracle_mod.compressstore(0, 64, vals, 64, words, 0, 64)
    assert out[:k].tolist() == [i for i in range(64) if bi["index"][i]]
    assert bi["index"] == bi["walk"]
    assert bi["index"] == np.unpackbits(np.array(bi["bytes"], np.uint8), bitorder="little").tolist()


def test_mask_uniform_ref_golden(oracle_mod):
    g = GOLD["mask_uniform_ref"]
    got = oracle_mod.mask_uniform_ref(len(g["bytes"]), g["selectivity"])
    assert got.tolist() == g["bytes"]
    # selectivity 0 selects nothing
    assert not oracle_mod.mask_uniform_ref(64, 0.0).any()


def test_mask_clusters_ref(oracle_mod):
    """create_bitmask (src/main.cu:59-108): k runs of floor(s*N)/k ones at
    i*N/k, bits MSB-first in each byte."""
    for sel, k, total in ((0.5, 1, 1024), (0.1, 8, 8192), (0.01, 8, 1 << 16), (0.9, 4, 4096)):
        got = oracle_mod.mask_clusters_ref(sel, k, total)
        bits = np.unpackbits(got, bitorder="big")
        size = int(np.float32(sel) * np.float32(total)) // k
        exp = np.zeros(total, np.uint8)
        for i in range(k):
            exp[i * (total // k):i * (total // k) + size] = 1
        assert np.array_equal(bits, exp)


# ---------------------------------------------------------------- formats

@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 7, 8, 11, 13, 16, 17, 21, 24, 31, 32, 33, 40, 47, 63, 64])
def test_formats_round_trip_and_closed_forms(oracle_mod, bits):
    rng = np.random.default_rng(bits)
    for n in (0, 1, 63, 64, 65, 1000):
        v = rand_vals(rng, n, bits)
        for fmt in range(5):
            if fmt == 4 and bits > 32:
                continue
            p = oracle_mod.pack(fmt, bits, v)
            assert len(p) == oracle_mod.packed_words(fmt, bits, n)
            assert np.array_equal(oracle_mod.unpack(fmt, bits, p, n), v)
        # slab: eps = floor(64/bits) values per word, dead bits on top are zero
        eps = 64 // bits
        s = oracle_mod.pack(1, bits, v)
        if n:
            assert len(s) == -(-n // eps)
            live = np.uint64(((1 << (eps * bits)) - 1) if eps * bits < 64 else 2**64 - 1)
            assert not (s & ~live).any()
        if bits in (8, 16, 32) and n:
            arr = s.view({8: np.uint8, 16: np.uint16, 32: np.uint32}[bits])[:n]
            assert np.array_equal(arr.astype(np.uint64), v)
        if bits == 64:
            assert np.array_equal(s, v)
        # bitstream == dense little-endian bit string
        b = oracle_mod.pack(2, bits, v)
        if n:
            allbits = np.unpackbits(b.view(np.uint8), bitorder="little")[:n * bits].reshape(n, bits)
            dec = (allbits.astype(np.uint64) << np.arange(bits, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
            assert np.array_equal(dec, v)
            tail = np.unpackbits(b.view(np.uint8), bitorder="little")[n * bits:]
            assert not tail.any()


def test_tight_containers(oracle_mod):
    for bits, dt in ((5, np.uint8), (12, np.uint16), (20, np.uint32), (33, np.uint64)):
        v = rand_vals(np.random.default_rng(bits), 77, bits)
        p = oracle_mod.pack(3, bits, v)
        assert np.array_equal(p.view(dt)[:77].astype(np.uint64), v)


# ---------------------------------------------------------------- operators

PRED_NP = {0: np.less, 1: np.less_equal, 2: np.greater, 3: np.greater_equal, 4: np.equal, 5: np.not_equal}


def test_filter_vs_numpy(oracle_mod):
    rng = np.random.default_rng(10)
    for bits in (1, 4, 11, 32, 64):
        n = 5000
        v = rand_vals(rng, n, bits)
        for fmt in (0, 1, 2):
            p = oracle_mod.pack(fmt, bits, v)
            for op, c in ((3, 1 << (bits - 1)), (0, 1024), (4, int(v[3])), (5, int(v[3]))):
                exp = lsb_mask_words(PRED_NP[op](v, np.uint64(c)))
                assert np.array_equal(oracle_mod.filter(fmt, bits, p, n, op, c), exp)


def test_compressstore_vs_numpy(oracle_mod):
    rng = np.random.default_rng(11)
    for bits in (1, 4, 11, 17, 32, 64):
        n = 3001
        v = rand_vals(rng, n, bits)
        for s in (0.0, 0.3, 1.0):
            sel = rng.random(n) < s
            m = lsb_mask_words(sel)
            for ofmt, ob in ((1, bits), (2, bits), (0, bits), (1, max(1, bits - 3))):
                k, out, ovf = oracle_mod.compressstore(1, bits, oracle_mod.pack(1, bits, v), n, m, ofmt, ob)
                exp = v[sel]
                assert k == sel.sum()
                assert ovf == int((exp & ~mask_of(ob)).astype(bool).sum())
                assert np.array_equal(oracle_mod.unpack(ofmt, ob, out, k), exp & mask_of(ob))
                # padding past k stays zero
                assert not out[oracle_mod.packed_words(ofmt, ob, k):].any()
            if s == 1.0:
                k, out, _ = oracle_mod.compressstore(1, bits, oracle_mod.pack(1, bits, v), n, m, 1, bits)
                assert np.array_equal(out, oracle_mod.pack(1, bits, v))  # identity


def test_golden_c1_small(oracle_mod):
    g = GOLD["c1_small"]
    n, bits = g["n"], g["bits"]
    words = np.array([int(x) for x in g["slab_words"]], dtype=np.uint64)
    vals = oracle_mod.prng(42, n).astype(np.uint64) & np.uint64(0x7FF)
    assert np.array_equal(oracle_mod.pack(1, bits, vals), words)
    k, out, mask = oracle_mod.cpu_filter_compress_u32(bits, words, n, 0, 1024)
    assert out.tolist() == g["out_u32"]
    assert [str(int(x)) for x in mask] == g["mask_words"]
    k2, out2, mask2 = oracle_mod.cpu_filter_compress_u32(bits, words, n, 0, 1024, threads=4)
    assert k2 == k and np.array_equal(out2, out) and np.array_equal(mask2, mask)


def np_binop(op, l, r, ob):
    with np.errstate(all="ignore"):
        if op == 0: v = l + r
        elif op == 1: v = l * r
        elif op == 2: v = np.where(r == 0, l, l % np.where(r == 0, 1, r))
        elif op == 3: v = np.where(r == 0, M64, l // np.where(r == 0, 1, r))
        elif op == 4:
            num = l * l + r * r + np.uint64(4) * l
            den = l * l + r * r + np.uint64(5) * l
            v = np.where(den == 0, M64, num // np.where(den == 0, 1, den))
        elif op == 5: v = ((l != 0) | (r != 0)).astype(np.uint64)
        elif op == 6: v = ((l != 0) ^ (r != 0)).astype(np.uint64)
        elif op == 7: v = l & r
        elif op == 8: v = l | r
        elif op == 9: v = l ^ r
        else: v = l - r
    return v & mask_of(ob)


def test_binary_vs_numpy(oracle_mod):
    rng = np.random.default_rng(12)
    for bits in (1, 4, 11, 32, 40, 64):
        n = 999
        a = rand_vals(rng, n, bits)
        b = rand_vals(rng, n, bits)
        b[::5] = 0
        a[::11] = 0
        for op in range(11):
            ob = oracle_mod.outbits(op, bits)
            for fmt, ofmt in ((1, 1), (2, 2), (0, 1)):
                out = oracle_mod.binary(op, fmt, bits, oracle_mod.pack(fmt, bits, a), oracle_mod.pack(fmt, bits, b), n,
                                        ofmt, ob)
                assert np.array_equal(oracle_mod.unpack(ofmt, ob, out, n), np_binop(op, a, b, ob)), (bits, op)


def test_outbits_rules(oracle_mod):
    assert [oracle_mod.outbits(0, w) for w in (1, 11, 63, 64)] == [2, 12, 64, 64]
    assert [oracle_mod.outbits(1, w) for w in (1, 11, 32, 40)] == [2, 22, 64, 64]
    assert [oracle_mod.outbits(2, w) for w in (1, 11, 64)] == [1, 11, 64]


def test_sum_vs_numpy(oracle_mod):
    rng = np.random.default_rng(13)
    for bits in (1, 11, 33, 64):
        v = rand_vals(rng, 10001, bits)
        for fmt in (0, 1, 2, 3):
            got = oracle_mod.sum(fmt, bits, oracle_mod.pack(fmt, bits, v), len(v))
            assert got == int(v.sum(dtype=np.uint64))  # numpy wraps mod 2^64 too
    # linearity: sum(a) + sum(b) == sum(a ++ b)
    a, b = rand_vals(rng, 500, 20), rand_vals(rng, 700, 20)
    s = lambda x: oracle_mod.sum(1, 20, oracle_mod.pack(1, 20, x), len(x))
    assert s(a) + s(b) == s(np.concatenate([a, b]))


def test_groupby_vs_bincount(oracle_mod):
    rng = np.random.default_rng(14)
    n = 20000
    keys = rng.integers(0, 1024, n).astype(np.uint64)
    vals = rand_vals(rng, n, 20)
    kp, vp = oracle_mod.pack(1, 10, keys), oracle_mod.pack(1, 20, vals)
    sums, counts, rows = oracle_mod.groupby(1, 10, kp, 1, 20, vp, n)
    assert rows == n
    assert np.array_equal(counts, np.bincount(keys.astype(np.int64), minlength=1024).astype(np.uint64))
    exp = np.zeros(1024, np.uint64)
    np.add.at(exp, keys.astype(np.int64), vals)
    assert np.array_equal(sums, exp)
    assert int(counts.sum()) == n and int(sums.sum(dtype=np.uint64)) == int(vals.sum(dtype=np.uint64))
    # chained filter v > 2^19
    sums2, counts2, rows2 = oracle_mod.groupby(1, 10, kp, 1, 20, vp, n, pred=(2, 1 << 19))
    sel = vals > (1 << 19)
    assert rows2 == sel.sum()
    assert np.array_equal(counts2, np.bincount(keys[sel].astype(np.int64), minlength=1024).astype(np.uint64))


def test_matmul_vs_numpy(oracle_mod):
    rng = np.random.default_rng(15)
    for (M, N, K) in ((1, 1, 1), (5, 7, 9), (17, 33, 65)):
        A = rng.integers(0, 128, (M, K)).astype(np.uint64)
        B = rng.integers(0, 128, (K, N)).astype(np.uint64)
        pa = np.concatenate([oracle_mod.pack(1, 7, A[i]) for i in range(M)])
        pb = np.concatenate([oracle_mod.pack(1, 7, B[i]) for i in range(K)])
        C = oracle_mod.matmul_packed(7, pa, pb, M, N, K)
        exp = A.astype(np.int64) @ B.astype(np.int64)
        assert np.array_equal(C, exp)
        # brute force on one element guards against a transposed operand
        assert C[M - 1, 0] == sum(int(A[M - 1, k]) * int(B[k, 0]) for k in range(K))


# ------------------------------------------------- this build's generators

def test_gen_mask_bernoulli_pinned(oracle_mod):
    n = 100003
    d = oracle_mod.prng(11, n).astype(np.uint64)
    # s chosen so the threshold equals one draw exactly: `<` leaves that bit clear
    s = float(d[17]) / 4294967296.0
    m = oracle_mod.gen_mask_bernoulli(11, n, s)
    bits = np.unpackbits(m.view(np.uint8), bitorder="little")[:n].astype(bool)
    assert not bits[17]
    assert np.array_equal(bits, d < d[17])
    assert np.array_equal(m, lsb_mask_words(d < d[17]))
    assert not oracle_mod.gen_mask_bernoulli(11, 1000, 0.0).any()
    full = oracle_mod.gen_mask_bernoulli(11, 1000, 1.0)
    assert np.array_equal(full, lsb_mask_words(np.ones(1000, bool)))  # bits past n stay zero
    for sel in (0.01, 0.5, 0.99):
        k = np.unpackbits(oracle_mod.gen_mask_bernoulli(3, n, sel).view(np.uint8)).sum()
        assert abs(k - sel * n) < 5 * np.sqrt(n * sel * (1 - sel)) + 1


def _runs(bits):
    d = np.diff(np.concatenate([[0], bits.astype(np.int8), [0]]))
    return np.nonzero(d == -1)[0] - np.nonzero(d == 1)[0]


def test_gen_mask_runs_pinned(oracle_mod):
    n = 1 << 22
    for sel in (0.01, 0.5, 0.99):
        m = oracle_mod.gen_mask_runs(12, n, sel, 256.0)
        bits = np.unpackbits(m.view(np.uint8), bitorder="little")[:n]
        assert abs(bits.mean() - sel) < 0.1 * max(sel, 0.05), sel
        # chunk c starts from draw c*4097: in state 1 iff that draw < round(s*2^32)
        chunks = n // 4096
        starts = np.array([oracle_mod.prng(12, 1, skip=c * 4097)[0] for c in range(0, chunks, 97)], np.uint64)
        assert np.array_equal(bits[np.arange(0, chunks, 97) * 4096].astype(bool),
                              starts < np.uint64(int(sel * 4294967296.0 + 0.5)))
        if sel == 0.5:
            r = _runs(bits)
            assert 150 < r.mean() < 300  # geometric, mean 256, cut at chunk ends
    # mean run 1: a 1-run never lasts past one bit inside a chunk
    bits = np.unpackbits(oracle_mod.gen_mask_runs(4, 1 << 16, 0.3, 1.0).view(np.uint8), bitorder="little")
    pairs = bits[:-1] & bits[1:]
    pairs[np.arange(4095, len(pairs), 4096)] = 0
    assert not pairs.any()


def test_gen_from_cdf_pinned(oracle_mod):
    n = 50000
    d = oracle_mod.prng(5, n)
    thr = np.unique(np.concatenate([d[:40], [4294967295]])).astype(np.uint32)
    v = oracle_mod.gen_from_cdf(5, thr, 16, n)
    exp = np.minimum(np.searchsorted(thr, d, side="right"), len(thr) - 1).astype(np.uint64) & np.uint64(0xFFFF)
    assert np.array_equal(v, exp)
    # element 0's draw is itself a threshold b: `<` maps it to b + 1
    b = int(np.searchsorted(thr, d[0]))
    assert thr[b] == d[0] and v[0] == b + 1
    # a Zipf(1.1) table over 1024 ranks: rank frequencies follow the weights
    p = 1.0 / np.arange(1, 1025) ** 1.1
    c = np.cumsum(p / p.sum())
    zt = np.minimum(np.floor(c * 4294967296.0), 4294967295).astype(np.uint64)
    zt[-1] = 4294967295
    z = oracle_mod.gen_from_cdf(42, zt.astype(np.uint32), 10, 1 << 20)
    h = np.bincount(z.astype(np.int64), minlength=1024) / (1 << 20)
    assert np.allclose(h[:8], (p / p.sum())[:8], rtol=0.05)


def test_gen_sorted_noise_pinned(oracle_mod):
    for bits, nb, n in ((3, 1, 1000), (11, 5, 40009), (31, 25, 40009), (40, 32, 10007)):
        v = oracle_mod.gen_sorted_noise(9, bits, nb, n)
        m = (1 << bits) - 1
        i = np.arange(n, dtype=np.float64)
        base = np.floor((i / float(n)) * float(m)).astype(np.uint64)
        noise = oracle_mod.prng(9, n).astype(np.uint64) & np.uint64((1 << nb) - 1)
        assert np.array_equal(v, np.minimum(base + noise, np.uint64(m)))
        assert (np.diff(base.astype(np.int64)) >= 0).all() and (v <= np.uint64(m)).all()
