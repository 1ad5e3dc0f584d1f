"""CPU oracle for the packed-column operators -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker. The product
(``amethyst_amd``) never imports it.

``oracle.c`` restates the reference's accessors and validation loops
(citations there); this module wraps it with ctypes/numpy. ``ref()`` returns
the optional library compiled from the reference's own header-only host code
(``oracle/_ref``), used to pin the restatement where it was built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

FMT_BASELINE, FMT_SLAB, FMT_BITSTREAM, FMT_TIGHT, FMT_U32 = 0, 1, 2, 3, 4
PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_EQ, PRED_NE = range(6)
(OP_ADD, OP_MUL, OP_MOD, OP_DIV, OP_MATH, OP_BOOL_SIMPLE, OP_BOOL_COMPLEX,
 OP_AND, OP_OR, OP_XOR, OP_SUB) = range(11)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile oracle.c (and oracle/_ref when the reference tree exists)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_prng_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, ctypes.c_size_t]
        L.or_gen_main_words.argtypes = [ctypes.c_uint64, _u64p, _u64p, ctypes.c_size_t]
        L.or_mask_uniform_ref.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double]
        L.or_mask_clusters_ref.argtypes = [_u8p, ctypes.c_float, ctypes.c_size_t, ctypes.c_size_t]
        L.or_gen_mask_bernoulli.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, _u64p]
        L.or_gen_mask_runs.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _u64p]
        L.or_gen_from_cdf.argtypes = [ctypes.c_uint64, _u32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, _u64p]
        L.or_gen_sorted_noise.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u64p]
        L.or_mask.argtypes = [ctypes.c_int]
        L.or_mask.restype = ctypes.c_uint64
        L.or_packed_words.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        L.or_packed_words.restype = ctypes.c_uint64
        L.or_get.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64]
        L.or_get.restype = ctypes.c_uint64
        L.or_pack.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64, _u64p]
        L.or_unpack.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64, _u64p]
        L.or_binop.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.or_binop.restype = ctypes.c_uint64
        L.or_outbits.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_outbits.restype = ctypes.c_int
        L.or_pred.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.or_pred.restype = ctypes.c_int
        L.or_filter.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_uint64, _u64p]
        L.or_compressstore.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64, _u64p,
                                       ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64, _u64p]
        L.or_compressstore.restype = ctypes.c_uint64
        L.or_binary.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p, _u64p, ctypes.c_uint64,
                                ctypes.c_int, ctypes.c_int, _u64p]
        L.or_sum.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64]
        L.or_sum.restype = ctypes.c_uint64
        L.or_groupby.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_int, _u64p,
                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u64p, _u64p]
        L.or_groupby.restype = ctypes.c_uint64
        L.or_matmul_packed.argtypes = [ctypes.c_int, _u64p, _u64p, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_uint64, _i64p]
        L.or_cpu_filter_compress_u32.argtypes = [ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_int,
                                                 ctypes.c_uint64, _u32p, _u64p]
        L.or_cpu_filter_compress_u32.restype = ctypes.c_uint64
        L.or_cpu_filter_compress_u32_mt.argtypes = [ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_int,
                                                    ctypes.c_uint64, _u32p, _u64p, ctypes.c_int]
        L.or_cpu_filter_compress_u32_mt.restype = ctypes.c_uint64
        L.or_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def ref():
    """The library built from the reference's own headers, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libamref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.amref_prng_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, ctypes.c_size_t]
        R.amref_bit_iterator_read.argtypes = [_u8p, ctypes.c_size_t, _u8p, _u8p]
        _REF = R
    return _REF


def _p(a, t):
    return a.ctypes.data_as(t)


def _u64(a):
    a = np.ascontiguousarray(a)
    assert a.dtype == np.uint64, a.dtype
    return a


# ------------------------------------------------------------------ generators

def prng(seed, n, skip=0):
    out = np.empty(n, dtype=np.uint32)
    lib().or_prng_fill(seed, skip, _p(out, _u32p), n)
    return out


def main_words(seed, n):
    a = np.empty(n, dtype=np.uint64)
    b = np.empty(n, dtype=np.uint64)
    lib().or_gen_main_words(seed, _p(a, _u64p), _p(b, _u64p), n)
    return a, b


def mask_uniform_ref(nbytes, selectivity):
    out = np.zeros(nbytes, dtype=np.uint8)
    lib().or_mask_uniform_ref(_p(out, _u8p), 0, nbytes, selectivity)
    return out


def mask_clusters_ref(selectivity, clusters, total):
    out = np.zeros(total // 8, dtype=np.uint8)
    lib().or_mask_clusters_ref(_p(out, _u8p), selectivity, clusters, total)
    return out


def gen_mask_bernoulli(seed, n, selectivity):
    out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    lib().or_gen_mask_bernoulli(seed, selectivity, n, _p(out, _u64p))
    return out[:(n + 63) // 64]


def gen_mask_runs(seed, n, selectivity, mean_run=256.0):
    out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    lib().or_gen_mask_runs(seed, selectivity, mean_run, n, _p(out, _u64p))
    return out[:(n + 63) // 64]


def gen_from_cdf(seed, thresholds, bits, n):
    thr = np.ascontiguousarray(thresholds, dtype=np.uint32)
    out = np.empty(max(1, n), dtype=np.uint64)
    lib().or_gen_from_cdf(seed, _p(thr, _u32p), len(thr), bits, n, _p(out, _u64p))
    return out[:n]


def gen_sorted_noise(seed, bits, noise_bits, n):
    out = np.empty(max(1, n), dtype=np.uint64)
    lib().or_gen_sorted_noise(seed, bits, noise_bits, n, _p(out, _u64p))
    return out[:n]


# --------------------------------------------------------------------- formats

def bitmask(bits):
    return int(lib().or_mask(bits))


def packed_words(fmt, bits, n):
    return int(lib().or_packed_words(fmt, bits, n))


def pack(fmt, bits, values):
    values = _u64(values)
    out = np.zeros(max(1, packed_words(fmt, bits, len(values))), dtype=np.uint64)
    lib().or_pack(fmt, bits, _p(values, _u64p), len(values), _p(out, _u64p))
    return out[:packed_words(fmt, bits, len(values))]


def unpack(fmt, bits, data, n):
    data = _u64(data)
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().or_unpack(fmt, bits, _p(data, _u64p), n, _p(out, _u64p))
    return out


def get(fmt, bits, data, i):
    data = _u64(data)
    return int(lib().or_get(fmt, bits, _p(data, _u64p), i))


# ------------------------------------------------------------------- operators

def binop(op, outbits, l, r):
    return int(lib().or_binop(op, outbits, l, r))


def outbits(op, bits):
    return int(lib().or_outbits(op, bits))


def filter(fmt, bits, data, n, pred_op, c):
    data = _u64(data)
    out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    lib().or_filter(fmt, bits, _p(data, _u64p), n, pred_op, c, _p(out, _u64p))
    return out[:(n + 63) // 64]


def compressstore(ifmt, ibits, data, n, mask, ofmt, obits, out_cap=None):
    """Returns (count, packed output of out_cap elements, overflow count)."""
    data = _u64(data)
    mask = _u64(mask)
    cap = n if out_cap is None else out_cap
    out = np.zeros(max(1, packed_words(ofmt, obits, cap)), dtype=np.uint64)
    ovf = np.zeros(1, dtype=np.uint64)
    k = lib().or_compressstore(ifmt, ibits, _p(data, _u64p), n, _p(mask, _u64p), ofmt, obits,
                               _p(out, _u64p), cap, _p(ovf, _u64p))
    return int(k), out[:packed_words(ofmt, obits, cap)], int(ovf[0])


def binary(op, fmt, bits, a, b, n, ofmt, obits):
    a = _u64(a)
    b = _u64(b)
    out = np.zeros(max(1, packed_words(ofmt, obits, n)), dtype=np.uint64)
    lib().or_binary(op, fmt, bits, _p(a, _u64p), _p(b, _u64p), n, ofmt, obits, _p(out, _u64p))
    return out[:packed_words(ofmt, obits, n)]


def sum(fmt, bits, data, n):
    data = _u64(data)
    return int(lib().or_sum(fmt, bits, _p(data, _u64p), n))


def groupby(kfmt, kbits, keys, vfmt, vbits, vals, n, pred=None):
    """Dense key-ordered (sums, counts, rows); pred = (op, c) or None."""
    keys = _u64(keys)
    vals = _u64(vals)
    g = 1 << kbits
    sums = np.zeros(g, dtype=np.uint64)
    counts = np.zeros(g, dtype=np.uint64)
    use = 0 if pred is None else 1
    pop, pc = (0, 0) if pred is None else pred
    rows = lib().or_groupby(kfmt, kbits, _p(keys, _u64p), vfmt, vbits, _p(vals, _u64p), n, use, pop, pc,
                            _p(sums, _u64p), _p(counts, _u64p))
    return sums, counts, int(rows)


def matmul_packed(bits, A, B, M, N, K):
    A = _u64(A)
    B = _u64(B)
    C = np.empty((M, N), dtype=np.int64)
    lib().or_matmul_packed(bits, _p(A, _u64p), _p(B, _u64p), M, N, K, _p(C, _i64p))
    return C


# --------------------------------------------------------- CPU baseline legs

def cpu_filter_compress_u32(bits, data, n, pred_op, c, threads=1):
    """Benchmark-1 path on the host: returns (count, u32 output, mask words)."""
    data = _u64(data)
    out = np.empty(max(1, n), dtype=np.uint32)
    mask = np.empty(max(1, (n + 63) // 64), dtype=np.uint64)
    if threads == 1:
        k = lib().or_cpu_filter_compress_u32(bits, _p(data, _u64p), n, pred_op, c, _p(out, _u32p),
                                             _p(mask, _u64p))
    else:
        k = lib().or_cpu_filter_compress_u32_mt(bits, _p(data, _u64p), n, pred_op, c, _p(out, _u32p),
                                                _p(mask, _u64p), threads)
    return int(k), out[:k], mask[:(n + 63) // 64]


def num_threads():
    return int(lib().or_num_threads())
