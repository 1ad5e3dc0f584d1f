"""MI355X (gfx950) packed-column operator engine -- Python host side.

A thin mirror of the C ABI in ``include/amethyst.h`` over ctypes. PyTorch is
used only for device memory and streams; every operator runs in the HIP
library ``amethyst_amd/libamethyst.so``. There is no CPU fallback: if the
library is missing, ``lib()`` raises.

The reference's operator surface is templated CUDA
(``kernel_<op>_<format><block_size, bits, [outbits], op>``, src/kernels/*,
host wrappers in src/lib.cuh:4-46). Here format, width and operator are
runtime arguments of one entry point per operator.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamethyst.so")
TUNE_TABLE = os.path.join(_HERE, "tune_table.txt")
_LIB = None

# formats (include/amethyst.h)
BASELINE, SLAB, BITSTREAM, TIGHT, U32 = 0, 1, 2, 3, 4
FORMAT_NAMES = {"baseline": BASELINE, "slab": SLAB, "bitstream": BITSTREAM, "tight": TIGHT, "u32": U32}
# predicates
LT, LE, GT, GE, EQ, NE = range(6)
# binary ops
(ADD, MUL, MOD, DIV, MATH, BOOL_SIMPLE, BOOL_COMPLEX, AND, OR, XOR, SUB) = range(11)
# group-by modes
GB_EXACT, GB_REF = 0, 1
# launch-shape knobs (am_knob)
(KNOB_FILTER_WORDS, KNOB_COMPRESS_CHUNKS, KNOB_BINARY_WORDS, KNOB_SUM_BLOCKS, KNOB_GROUPBY_BLOCKS,
 KNOB_PIPELINE_TILE) = range(6)

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported configuration", 3: "HIP runtime error",
          4: "workspace missing or too small", 5: "output capacity too small",
          6: "look-back wait timed out (workspace poisoned until re-initialised)",
          7: "input exceeds the no-allocation entry point's fixed scratch"}


class AmethystError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {STATUS.get(code, code)} (status {code})")
        self.code = code


class am_column(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_uint64), ("bits", ctypes.c_uint32),
                ("format", ctypes.c_uint32)]


class am_matmul_desc(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("C_packed", ctypes.c_void_p),
                ("C_i32", ctypes.c_void_p), ("M", ctypes.c_uint32), ("N", ctypes.c_uint32), ("K", ctypes.c_uint32),
                ("out_bits", ctypes.c_uint32)]


class am_predicate(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("c", ctypes.c_uint64)]


class am_pipeline_col(ctypes.Structure):
    _fields_ = [("a", am_column), ("b", am_column), ("s", am_column), ("pred", am_predicate),
                ("stats", ctypes.c_void_p), ("variant", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def build(verbose=False):
    """Compile the HIP library in-tree for gfx950 (hipcc via make)."""
    import subprocess
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), "-j8"]
    if not verbose:
        cmd.insert(1, "-s")
    subprocess.check_call(cmd)


def lib():
    """Load libamethyst.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run amethyst_amd.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        colp = ctypes.POINTER(am_column)
        predp = ctypes.POINTER(am_predicate)
        sigs = {
            "am_abi_version": ([], i32),
            "am_status_string": ([i32], ctypes.c_char_p),
            "am_packed_words": ([u32, u32, u64], u64),
            "am_outbits": ([u32, u32], u32),
            "am_pred_width": ([am_predicate, u32], u32),
            "am_gen_bitmask": ([u32], u64),
            "am_workspace_bytes": ([u64], u64),
            "am_workspace_init": ([vp, u64, vp], i32),
            "am_workspace_status": ([vp], i32),
            "am_read_device": ([vp, vp, u64, vp], i32),
            "am_prng_fill_u32": ([u64, u64, vp, u64, vp], i32),
            "am_gen_uniform": ([u64, colp, vp], i32),
            "am_gen_from_cdf": ([u64, vp, u32, colp, vp], i32),
            "am_gen_mask_bernoulli": ([u64, ctypes.c_double, vp, u64, vp], i32),
            "am_gen_mask_runs": ([u64, ctypes.c_double, ctypes.c_double, vp, u64, vp], i32),
            "am_gen_sorted_noise": ([u64, u32, colp, vp], i32),
            "am_gen_mask_uniform_ref": ([u64, ctypes.c_double, vp, u64, u64, vp], i32),
            "am_gen_mask_clusters_ref": ([ctypes.c_float, u64, u64, vp, vp], i32),
            "am_stream_copy": ([vp, vp, u64, vp], i32),
            "am_convert": ([colp, colp, vp], i32),
            "am_filter": ([colp, am_predicate, vp, vp], i32),
            "am_compressstore": ([colp, vp, predp, colp, vp, vp, vp, u64, vp], i32),
            "am_compressstore_noalloc": ([colp, vp, predp, colp, vp, vp, vp], i32),
            "am_binary": ([u32, colp, colp, colp, vp], i32),
            "am_sum": ([colp, vp, vp, u64, vp], i32),
            "am_groupby": ([colp, colp, predp, u32, vp, vp, vp, vp, vp, u64, vp], i32),
            "am_groupby_workspace_bytes": ([u64, u32, u32], u64),
            "am_pipeline_workspace_bytes": ([vp, u32], u64),
            "am_pipeline_prepare": ([vp, u32, vp, u64, vp, vp], i32),
            "am_pipeline_run": ([vp, u64, u32, vp], i32),
            "am_plan_scratch_bytes": ([u32], u64),
            "am_plan_features": ([vp, u32, vp, vp, vp], i32),
            "am_plan_infer": ([vp, u32, vp, u32, vp, vp, ctypes.c_float, vp, vp, vp, vp], i32),
            "am_matmul_out_bits": ([u32, u32], u32),
            "am_matmul_workspace_bytes": ([vp, vp, vp, u32], u64),
            "am_matmul_packed": ([u32, vp, vp, u32, u32, u32, u32, vp, vp, vp, u64, vp], i32),
            "am_matmul_packed_batched": ([u32, vp, u32, vp, u64, vp], i32),
            "am_matmul_plan_create": ([u32, vp, u32, vp, u64, vp, vp], i32),
            "am_matmul_plan_run": ([vp, vp], i32),
            "am_matmul_plan_destroy": ([vp], None),
            "am_tune_set": ([u32, u32, u32, u32, u32], i32),
            "am_tune_get": ([u32, u32, u32, u32], u32),
            "am_tune_clear": ([], None),
            "am_tune_load": ([ctypes.c_char_p], i32),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
        # the autotuned launch shapes (amethyst_amd/autotune.py) unless
        # AMETHYST_TUNE=0 asks for the built-in ones
        if os.environ.get("AMETHYST_TUNE", "1") != "0" and os.path.exists(TUNE_TABLE):
            if L.am_tune_load(TUNE_TABLE.encode()) < 0:
                raise RuntimeError(f"malformed launch-shape table {TUNE_TABLE}")
    return _LIB


def exported_symbols():
    """Names of the C entry points this module binds."""
    lib()
    return [n for n in dir(_LIB) if n.startswith("am_")]


def _check(code, what):
    if code != 0:
        raise AmethystError(code, what)


# ----------------------------------------------------------------- pure helpers

def packed_words(fmt, bits, n):
    return int(lib().am_packed_words(fmt, bits, n))


def outbits(op, bits):
    return int(lib().am_outbits(op, bits))


def pred_width(op, c, bits):
    return int(lib().am_pred_width(am_predicate(op, 0, c), bits))


def tune_set(knob, bits, in_fmt, out_fmt, value):
    _check(lib().am_tune_set(knob, bits, in_fmt, out_fmt, value), "am_tune_set")


def tune_get(knob, bits, in_fmt, out_fmt=0):
    return int(lib().am_tune_get(knob, bits, in_fmt, out_fmt))


def tune_clear():
    lib().am_tune_clear()


def tune_load(path):
    got = int(lib().am_tune_load(path.encode()))
    if got < 0:
        raise RuntimeError(f"cannot read launch-shape table {path}")
    return got


def workspace_bytes(max_elements):
    return int(lib().am_workspace_bytes(max_elements))


# ------------------------------------------------------------- device columns

def _torch():
    import torch
    return torch


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def device_words(words, device="cuda"):
    """A zero-filled device buffer of `words` uint64 words (int64 storage)."""
    torch = _torch()
    return torch.zeros(max(1, int(words)), dtype=torch.int64, device=device)


class Column:
    """`n` values of `bits` bits in `fmt`, held in a device buffer of uint64 words."""

    def __init__(self, n, bits, fmt, buf=None, device="cuda"):
        self.n, self.bits, self.fmt = int(n), int(bits), int(fmt)
        w = packed_words(self.fmt, self.bits, self.n)
        if w == 0 and self.n:
            raise ValueError(f"invalid format/width: fmt={fmt} bits={bits}")
        self.buf = buf if buf is not None else device_words(w, device)
        if self.buf.numel() < max(1, w):
            raise ValueError("buffer too small for the column")

    @property
    def words(self):
        return packed_words(self.fmt, self.bits, self.n)

    def c(self):
        return am_column(self.buf.data_ptr(), self.n, self.bits, self.fmt)

    def host_words(self):
        """Packed words copied to the host as a numpy uint64 array."""
        return self.buf[:self.words].cpu().numpy().view(np.uint64).copy()

    @classmethod
    def from_host_words(cls, words, n, bits, fmt, device="cuda"):
        torch = _torch()
        w = np.ascontiguousarray(words, dtype=np.uint64)
        need = packed_words(fmt, bits, n)
        buf = torch.zeros(max(1, need), dtype=torch.int64, device=device)
        if need:
            buf[:need].copy_(torch.from_numpy(w[:need].view(np.int64)))
        return cls(n, bits, fmt, buf=buf)


class Workspace:
    """Zero-initialised workspace for the single-pass operators."""

    def __init__(self, max_elements, device="cuda"):
        self.bytes = workspace_bytes(max_elements)
        self.buf = device_words((self.bytes + 7) // 8, device)  # torch.zeros: already zero

    @classmethod
    def for_groupby(cls, n, key_bits, value_bits, device="cuda"):
        """A workspace large enough for am_groupby over n rows at these widths."""
        ws = cls.__new__(cls)
        ws.bytes = max(workspace_bytes(0), int(lib().am_groupby_workspace_bytes(n, key_bits, value_bits)))
        ws.buf = device_words((ws.bytes + 7) // 8, device)
        return ws

    def ptr(self):
        return ctypes.c_void_p(self.buf.data_ptr())

    def check(self):
        """Raise if a single-pass launch on this workspace gave up waiting
        (bounded look-back wait); synchronous."""
        _check(lib().am_workspace_status(self.ptr()), "am_workspace_status")


# ----------------------------------------------------------------- operators

def prng_fill(seed, n, offset=0, stream=None):
    torch = _torch()
    out = torch.empty(max(1, n), dtype=torch.int32, device="cuda")
    _check(lib().am_prng_fill_u32(seed, offset, ctypes.c_void_p(out.data_ptr()), n, _stream(stream)),
           "am_prng_fill_u32")
    return out[:n]


def gen_uniform(seed, n, bits, fmt, stream=None):
    col = Column(n, bits, fmt)
    c = col.c()
    _check(lib().am_gen_uniform(seed, ctypes.byref(c), _stream(stream)), "am_gen_uniform")
    return col


def gen_from_cdf(seed, thresholds, n, bits, fmt, stream=None):
    """Values drawn by inverse CDF: value = first b with draw < thresholds[b]."""
    torch = _torch()
    thr = torch.as_tensor(np.ascontiguousarray(thresholds, dtype=np.uint32).view(np.int32)).cuda()
    col = Column(n, bits, fmt)
    c = col.c()
    _check(lib().am_gen_from_cdf(seed, ctypes.c_void_p(thr.data_ptr()), len(thresholds), ctypes.byref(c),
                                 _stream(stream)), "am_gen_from_cdf")
    torch.cuda.current_stream().synchronize()
    return col


def gen_mask_bernoulli(seed, n, selectivity, stream=None):
    m = device_words((n + 63) // 64)
    _check(lib().am_gen_mask_bernoulli(seed, selectivity, ctypes.c_void_p(m.data_ptr()), n, _stream(stream)),
           "am_gen_mask_bernoulli")
    return m


def gen_sorted_noise(seed, n, bits, fmt, noise_bits, stream=None):
    """Ascending values with noise (the uneven batch's sorted columns)."""
    col = Column(n, bits, fmt)
    c = col.c()
    _check(lib().am_gen_sorted_noise(seed, noise_bits, ctypes.byref(c), _stream(stream)), "am_gen_sorted_noise")
    return col


def gen_mask_uniform_ref(nbytes, selectivity, seed=42, stream=None):
    """The reference's generate_mask_uniform bytes (MSB-first within a byte)
    as a device buffer of ceil(nbytes/8) words."""
    m = device_words((nbytes + 7) // 8)
    _check(lib().am_gen_mask_uniform_ref(seed, selectivity, ctypes.c_void_p(m.data_ptr()), 0, nbytes,
                                         _stream(stream)), "am_gen_mask_uniform_ref")
    return m


def gen_mask_clusters_ref(selectivity, clusters, total, stream=None):
    """The reference's create_bitmask bytes (total/8 of them) as a device
    buffer of ceil(total/64) words (bytes past total/8 zero)."""
    m = device_words((total + 63) // 64)
    _check(lib().am_gen_mask_clusters_ref(selectivity, clusters, total, ctypes.c_void_p(m.data_ptr()),
                                          _stream(stream)), "am_gen_mask_clusters_ref")
    return m


def stream_copy(src, dst, stream=None):
    """Copy device int64 tensor `src` into `dst` (same length, 16-byte aligned)
    with non-temporal 16-byte loads and stores (the sustained copy rate)."""
    assert src.numel() == dst.numel()
    _check(lib().am_stream_copy(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), src.numel(),
                                _stream(stream)), "am_stream_copy")
    return dst


def gen_mask_runs(seed, n, selectivity, mean_run=256.0, stream=None):
    m = device_words((n + 63) // 64)
    _check(lib().am_gen_mask_runs(seed, selectivity, mean_run, ctypes.c_void_p(m.data_ptr()), n, _stream(stream)),
           "am_gen_mask_runs")
    return m


def convert(col, bits, fmt, stream=None):
    out = Column(col.n, bits, fmt)
    a, o = col.c(), out.c()
    _check(lib().am_convert(ctypes.byref(a), ctypes.byref(o), _stream(stream)), "am_convert")
    return out


def filter(col, op, c, mask_out=None, stream=None):
    """Predicate filter -> LSB-first bitmask (device int64 buffer of ceil(n/64) words)."""
    m = mask_out if mask_out is not None else device_words((col.n + 63) // 64)
    a = col.c()
    _check(lib().am_filter(ctypes.byref(a), am_predicate(op, 0, c), ctypes.c_void_p(m.data_ptr()), _stream(stream)),
           "am_filter")
    return m


def compressstore(col, out_bits, out_fmt, mask=None, pred=None, want_mask=False, capacity=None,
                  ws=None, out=None, stats=None, mask_out=None, stream=None):
    """Compaction; returns (out Column (n = capacity), stats device tensor, mask or None).

    stats holds {selected count, values that did not fit out_bits}.
    """
    if (mask is None) == (pred is None):
        raise ValueError("give exactly one of mask / pred")
    cap = col.n if capacity is None else int(capacity)
    out = out if out is not None else Column(cap, out_bits, out_fmt)
    ws = ws if ws is not None else Workspace(col.n)
    stats = stats if stats is not None else device_words(2)
    if want_mask and mask_out is None:
        mask_out = device_words((col.n + 63) // 64)
    a, o = col.c(), out.c()
    pp = ctypes.byref(am_predicate(pred[0], 0, pred[1])) if pred is not None else None
    mptr = ctypes.c_void_p(mask.data_ptr()) if mask is not None else None
    moptr = ctypes.c_void_p(mask_out.data_ptr()) if mask_out is not None else None
    _check(lib().am_compressstore(ctypes.byref(a), mptr, pp, ctypes.byref(o), moptr,
                                  ctypes.c_void_p(stats.data_ptr()), ws.ptr(), ws.bytes, _stream(stream)),
           "am_compressstore")
    return out, stats, mask_out


def binary(op, a, b, out_bits=None, out_fmt=None, out=None, stream=None):
    ob = outbits(op, a.bits) if out_bits is None else out_bits
    of = a.fmt if out_fmt is None else out_fmt
    out = out if out is not None else Column(a.n, ob, of)
    ca, cb, co = a.c(), b.c(), out.c()
    _check(lib().am_binary(op, ctypes.byref(ca), ctypes.byref(cb), ctypes.byref(co), _stream(stream)), "am_binary")
    return out


def sum(col, ws=None, out=None, stream=None):
    """Wrapping uint64 sum; returns a 1-word device tensor."""
    ws = ws if ws is not None else Workspace(0)
    out = out if out is not None else device_words(1)
    a = col.c()
    _check(lib().am_sum(ctypes.byref(a), ctypes.c_void_p(out.data_ptr()), ws.ptr(), ws.bytes, _stream(stream)),
           "am_sum")
    return out


def groupby(keys, values, mode=GB_EXACT, value_pred=None, want_counts=True, ws=None, stream=None):
    """Key-ordered group-by; returns (out_keys, out_sums, out_counts, result) device tensors.

    result = {groups, sum bits, count bits, rows aggregated}; the outputs are
    slab-packed at keys.bits / result[1] / result[2].
    """
    g = 1 << keys.bits
    cap = packed_words(SLAB, 1, g)  # enough for any width (1 bit: 64 per word ... 64 bits: g words)
    cap = max(cap, g)
    ok, osum, ocnt = device_words(cap), device_words(cap), device_words(cap) if want_counts else None
    res = device_words(4)
    ws = ws if ws is not None else Workspace.for_groupby(keys.n, keys.bits, values.bits)
    ck, cv = keys.c(), values.c()
    pp = ctypes.byref(am_predicate(value_pred[0], 0, value_pred[1])) if value_pred is not None else None
    _check(lib().am_groupby(ctypes.byref(ck), ctypes.byref(cv), pp, mode, ctypes.c_void_p(ok.data_ptr()),
                            ctypes.c_void_p(osum.data_ptr()),
                            ctypes.c_void_p(ocnt.data_ptr()) if ocnt is not None else None,
                            ctypes.c_void_p(res.data_ptr()), ws.ptr(), ws.bytes, _stream(stream)), "am_groupby")
    return ok, osum, ocnt, res


class PipelineBatch:
    """A prepared batch of fused filter -> compressstore(a+b) -> sum chains.

    columns: list of (a Column, b Column, (pred_op, c), s_fmt, s_bits).
    variant: elements per lane per tile for every column (8, 16 or 32, 0 = the
    launch-shape table's choice), or a list with one per column.
    run() is one launch over every tile of every column (capturable in a
    HIP graph); results: per column the output Column s (capacity n) and
    stats {count, overflow, total, 0}.
    """

    def __init__(self, columns, variant=32, outs=None, stream=None):
        torch = _torch()
        n = len(columns)
        self.cols = columns
        variants = list(variant) if isinstance(variant, (list, tuple, np.ndarray)) else [variant] * n
        self.outs = []
        self.stats = device_words(4 * n)
        arr = (am_pipeline_col * n)()
        ns = (ctypes.c_uint64 * n)()
        for i, (a, b, pred, sfmt, sbits) in enumerate(columns):
            out = outs[i] if outs is not None else Column(a.n, sbits, sfmt)
            self.outs.append(out)
            arr[i] = am_pipeline_col(a.c(), b.c(), out.c(), am_predicate(pred[0], 0, pred[1]),
                                     self.stats.data_ptr() + 32 * i, int(variants[i]), 0)
            ns[i] = a.n
        self.arr = arr
        self.ws_bytes = int(lib().am_pipeline_workspace_bytes(ns, n))
        self.ws = device_words((self.ws_bytes + 7) // 8)
        tt = ctypes.c_uint32(0)
        _check(lib().am_pipeline_prepare(arr, n, ctypes.c_void_p(self.ws.data_ptr()), self.ws_bytes,
                                         ctypes.byref(tt), _stream(stream)), "am_pipeline_prepare")
        self.total_tiles = tt.value
        self.n = n

    def run(self, stream=None):
        _check(lib().am_pipeline_run(ctypes.c_void_p(self.ws.data_ptr()), self.ws_bytes, self.total_tiles,
                                     _stream(stream)), "am_pipeline_run")

    def capture(self):
        """A HIP graph holding one run (torch.cuda.CUDAGraph; replay() to run).
        The batch is run once eagerly on a side stream first."""
        torch = _torch()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g

    def results(self):
        st = u64(self.stats).reshape(self.n, 4)
        return self.outs, st


def matmul_out_bits(bits, K):
    return int(lib().am_matmul_out_bits(bits, K))


def matrix_words(rows, cols, bits):
    """Words of a row-padded slab matrix (each row padded to whole slabs)."""
    eps = 64 // bits
    return rows * ((cols + eps - 1) // eps)


def matmul_packed_batched(bits, problems, want_i32=False, out_bits=None, stream=None):
    """problems: list of (A_words_tensor, B_words_tensor, M, N, K).

    Returns a list of (C_packed tensor, C_i32 tensor or None, out_bits).
    Every word of every output is written by the library, so the outputs and
    the workspace are left uninitialised (one allocation each)."""
    torch = _torch()
    n = len(problems)
    Ms = (ctypes.c_uint32 * n)(*[p[2] for p in problems])
    Ns = (ctypes.c_uint32 * n)(*[p[3] for p in problems])
    Ks = (ctypes.c_uint32 * n)(*[p[4] for p in problems])
    wsb = int(lib().am_matmul_workspace_bytes(Ms, Ns, Ks, n))
    ws = torch.empty((wsb + 7) // 8, dtype=torch.int64, device="cuda")
    obs = [out_bits if out_bits else matmul_out_bits(bits, p[4]) for p in problems]
    pw = [matrix_words(M, N, ob) for (_, _, M, N, _), ob in zip(problems, obs)]
    tot_p = tot_c = 0  # (the module's own sum() shadows the builtin)
    for (_, _, M, N, _), w in zip(problems, pw):
        tot_p += w
        tot_c += M * N
    cp_all = torch.empty(max(1, tot_p), dtype=torch.int64, device="cuda")
    ci_all = torch.empty(max(1, tot_c), dtype=torch.int32, device="cuda") if want_i32 else None
    descs = (am_matmul_desc * n)()
    outs = []
    po = co = 0
    for i, (A, B, M, N, K) in enumerate(problems):
        cp = cp_all[po:po + pw[i]]
        ci = ci_all[co:co + M * N] if want_i32 else None
        po += pw[i]
        co += M * N
        descs[i] = am_matmul_desc(A.data_ptr(), B.data_ptr(), cp.data_ptr(), ci.data_ptr() if ci is not None else None,
                                  M, N, K, obs[i])
        outs.append((cp, ci, obs[i]))
    _check(lib().am_matmul_packed_batched(bits, descs, n, ctypes.c_void_p(ws.data_ptr()), wsb, _stream(stream)),
           "am_matmul_packed_batched")
    outs.append(ws)  # keep the workspace alive until the caller synchronises
    return outs


class MatmulPlan:
    """A prepared batch of packed products (am_matmul_plan_*): the header is
    uploaded once; run() launches the kernels only (graph-capturable).
    problems as for matmul_packed_batched; outs like its return value."""

    def __init__(self, bits, problems, want_i32=False, out_bits=None, stream=None):
        torch = _torch()
        n = len(problems)
        Ms = (ctypes.c_uint32 * n)(*[p[2] for p in problems])
        Ns = (ctypes.c_uint32 * n)(*[p[3] for p in problems])
        Ks = (ctypes.c_uint32 * n)(*[p[4] for p in problems])
        self.ws_bytes = int(lib().am_matmul_workspace_bytes(Ms, Ns, Ks, n))
        self.ws = torch.empty((self.ws_bytes + 7) // 8, dtype=torch.int64, device="cuda")
        self.outs, descs = [], (am_matmul_desc * n)()
        for i, (A, B, M, N, K) in enumerate(problems):
            ob = out_bits if out_bits else matmul_out_bits(bits, K)
            cp = torch.empty(max(1, matrix_words(M, N, ob)), dtype=torch.int64, device="cuda")
            ci = torch.empty(M * N, dtype=torch.int32, device="cuda") if want_i32 else None
            descs[i] = am_matmul_desc(A.data_ptr(), B.data_ptr(), cp.data_ptr(), ci.data_ptr() if ci is not None else None,
                                      M, N, K, ob)
            self.outs.append((cp, ci, ob))
        self.problems = problems  # keep the inputs alive
        self.handle = ctypes.c_void_p()
        _check(lib().am_matmul_plan_create(bits, descs, n, ctypes.c_void_p(self.ws.data_ptr()), self.ws_bytes,
                                           ctypes.byref(self.handle), _stream(stream)), "am_matmul_plan_create")

    def run(self, stream=None):
        _check(lib().am_matmul_plan_run(self.handle, _stream(stream)), "am_matmul_plan_run")

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value and _LIB is not None:
            _LIB.am_matmul_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()


def u64(t):
    """Device int64 tensor -> host numpy uint64 array."""
    return t.cpu().numpy().view(np.uint64)
