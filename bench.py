"""Benchmark: achieved HBM GB/s of the packed-column operators on one MI355X.

Headline step (the JSON line's `value`): one fused filter + compressstore over
a 1 GB-equivalent column -- N = 2^28 uint32 values uniform in [0, 2^11)
(PCG32 seed 42, generated on the GPU), slab-packed at 11 bits, predicate
v < 1024 (~50 % selectivity), output uncompressed uint32 + LSB-first bitmask.
value = algorithmic bytes / time, bytes = packed input + 4*K + N/8.

The rest of the operator sweep (widths {4,8,11,17,24,32}, output formats,
selectivities, mask shapes, binary add/and, u64 sum, group-by) is measured in
the same run and reported under "ops" (median of repeated launches, HIP
events on the launch stream).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0
N_FULL = 1 << 28
WIDTHS = (4, 8, 11, 17, 24, 32)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--quick", action="store_true", help="headline only, no sweep")
    p.add_argument("--legs", default="sweep,groupby,matmul,batched,libbase",
                   help="comma-separated legs to run after the headline (sweep, groupby, matmul, batched, "
                        "libbase: hipCUB select/reduce over u64 elements beside the library's operators)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--reps", type=int, default=10, help="launches per sweep point")
    p.add_argument("--csv", default=None, metavar="DIR",
                   help="also write every timed launch of the sweep in the reference's CSV schema under DIR")
    p.add_argument("--case", default=None, choices=sorted(CASES),
                   help="run one workload --steps times and exit (for rocprofv3; prints its algorithmic bytes)")
    p.add_argument("--profile-copy", action="store_true",
                   help="also run a plain 1 GiB device copy (reference point for HBM counters)")
    return p.parse_args()


CASES = ("headline", "compress_slab_w4", "compress_slab_w11", "pipeline", "copy_1gib", "groupby_u24", "groupby_z16",
         "add_slab_w4", "add_bitstream_w11", "sum_bitstream_w11")


def gbytes(b, ms):
    return b / (ms * 1e-3) / 1e9


class Timer:
    """Times launches on torch's current stream (the stream the ops run on)."""

    def __init__(self, torch):
        self.torch = torch
        self.last = []

    def median_ms(self, fn, reps, warm=2):
        torch = self.torch
        for _ in range(warm):
            fn()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in evs:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        self.last = [a.elapsed_time(b) for a, b in evs]
        return statistics.median(self.last)


def read_traffic(name):
    """FETCH/WRITE KiB per dispatch from a committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line:
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    if "fetch_kib_per_dispatch" not in vals or "write_kib_per_dispatch" not in vals:
        return None
    # each counter scaled by what it reads for a plain copy kernel of known size
    fs = vals.get("fetch_scale", 1.0)
    ws = vals.get("write_scale", 1.0)
    return (vals["fetch_kib_per_dispatch"] * fs + vals["write_kib_per_dispatch"] * ws) * 1024.0


def cpu_baseline(am, torch, col, pred_c):
    """The oracle's restatement of the reference's host loop, timed on the host
    on a bounded sample: the first 2^24 values of the same column."""
    import oracle
    nsamp = 1 << 24
    bits = col.bits
    words = col.buf[:am.packed_words(am.SLAB, bits, nsamp)].cpu().numpy().view(np.uint64).copy()
    k = None
    res = {}
    for threads in (1, oracle.num_threads()):
        times = []
        for rep in range(4):
            t0 = time.perf_counter()
            k, out, mask = oracle.cpu_filter_compress_u32(bits, words, nsamp, am.LT, pred_c, threads=threads)
            times.append(time.perf_counter() - t0)
        t = statistics.median(times[1:])
        b = am.packed_words(am.SLAB, bits, nsamp) * 8 + 4 * k + (nsamp + 63) // 64 * 8
        res[threads] = (b / t / 1e9, t)
    st = res[1]
    mt_threads = max(res)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(st[0], 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(), "threads_parallel": mt_threads,
            "sample": f"first 2^24 of the 2^28 w=11 slab column (filter v<1024 + compressstore to u32 + mask), "
                      f"single-thread faithful per-element loop, median of 3 after 1 warm-up ({st[1]*1e3:.1f} ms)",
            "parallel": {"value": round(res[mt_threads][0], 3), "cores": mt_threads,
                         "ms": round(res[mt_threads][1] * 1e3, 2)}}


def needed_line_bytes(torch, mask, n, eps):
    """Bytes of the 128-byte input lines (16 slab words = 16*eps elements)
    that hold at least one selected element: the input a compaction that
    skips empty regions has to read."""
    words = mask[:(n + 63) // 64]
    bits = ((words.view(-1, 1) >> torch.arange(64, device=words.device)) & 1).view(-1)[:n].bool()
    L = 16 * eps
    pad = (-n) % L
    if pad:
        bits = torch.cat([bits, torch.zeros(pad, dtype=torch.bool, device=bits.device)])
    return int(bits.view(-1, L).any(dim=1).sum().item()) * 128


def sweep(am, torch, timer, reps):
    """Per-op achieved GB/s at N = 2^28 (algorithmic bytes)."""
    ops = []
    n = N_FULL
    ws = am.Workspace(n)
    stats = am.device_words(2)
    mask_words = (n + 63) // 64

    def rec(d, ms, byt, fn=None, key=None):
        d.update({"n": n, "ms": ms, "GBs": gbytes(byt, ms), "times": list(timer.last)})
        tv = am.tune_get(*key) if key else 0
        if tv:  # a tuned launch shape: time the built-in one beside it
            am.tune_set(*key, 0)
            d["ms_default"] = round(timer.median_ms(fn, reps), 4)
            am.tune_set(*key, tv)
            d["tuned"] = {"knob": key[0], "value": tv}
        ops.append(d)

    for w in WIDTHS:
        print(f"sweep w={w}", file=sys.stderr, flush=True)
        a = am.gen_uniform(42, n, w, am.SLAB)
        b = am.gen_uniform(43, n, w, am.SLAB)
        in_b = a.words * 8
        eps = 64 // w
        torch.cuda.synchronize()

        # filter -> mask
        mo = am.device_words(mask_words)
        f = lambda: am.filter(a, am.LT, 1 << (w - 1), mask_out=mo)
        ms = timer.median_ms(f, reps)
        rec({"op": "filter", "w": w, "in": "slab", "out": "mask"}, ms, in_b + mask_words * 8, f,
            (am.KNOB_FILTER_WORDS, w, am.SLAB, 0))

        # fused filter + compressstore at 1/50/99 % selectivity, three output formats
        for frac in (0.01, 0.5, 0.99):
            c = max(1, int(round(frac * (1 << w))))
            pw = am.pred_width(am.LT, c, w)
            for fo, ob, label in ((am.U32, w, "u32"), (am.SLAB, w, "slab_same_w"), (am.SLAB, pw, "slab_min_w")):
                out = am.Column(n, ob, fo)
                am.compressstore(a, ob, fo, pred=(am.LT, c), ws=ws, out=out, stats=stats)
                torch.cuda.synchronize()
                k = int(am.u64(stats)[0])
                f = lambda: am.compressstore(a, ob, fo, pred=(am.LT, c), ws=ws, out=out, stats=stats)
                ms = timer.median_ms(f, reps)
                rec({"op": "filter+compressstore", "w": w, "sel": frac, "in": "slab", "out": label, "ob": ob},
                    ms, in_b + am.packed_words(fo, ob, k) * 8, f, (am.KNOB_COMPRESS_CHUNKS, w, am.SLAB, fo))
                del out

        # mask-driven compressstore (reference semantics: slab -> slab same w):
        # uniform Bernoulli masks, clustered masks (alternating geometric
        # runs, mean 256 ones), and the reference driver's own create_bitmask
        # masks with 1 and 8 equal clusters
        shapes = (("uniform", (0.01, 0.5, 0.99)), ("clustered", (0.01, 0.5, 0.99)),
                  ("ref_1_cluster", (0.01, 0.5)), ("ref_8_clusters", (0.01, 0.5)))
        for shape, fracs in shapes:
            for frac in fracs:
                if shape == "uniform":
                    m = am.gen_mask_bernoulli(7, n, frac)
                elif shape == "clustered":
                    m = am.gen_mask_runs(7, n, frac, 256.0)
                else:
                    m = am.gen_mask_clusters_ref(frac, 1 if shape == "ref_1_cluster" else 8, n)
                out = am.Column(n, w, am.SLAB)
                am.compressstore(a, w, am.SLAB, mask=m, ws=ws, out=out, stats=stats)
                torch.cuda.synchronize()
                k = int(am.u64(stats)[0])
                f = lambda: am.compressstore(a, w, am.SLAB, mask=m, ws=ws, out=out, stats=stats)
                ms = timer.median_ms(f, reps)
                out_b = am.packed_words(am.SLAB, w, k) * 8
                need = needed_line_bytes(torch, m, n, eps)
                rec({"op": "compressstore(mask)", "w": w, "sel": frac, "mask": shape, "in": "slab",
                     "out": "slab_same_w", "needed_GBs": round(gbytes(need + mask_words * 8 + out_b, ms), 1)},
                    ms, in_b + mask_words * 8 + out_b, f, (am.KNOB_COMPRESS_CHUNKS, w, am.SLAB, am.SLAB))
                del out, m

        # binary add (w+1) and and (w)
        for op, name in ((am.ADD, "add"), (am.AND, "and")):
            ob = am.outbits(op, w)
            out = am.Column(n, ob, am.SLAB)
            f = lambda: am.binary(op, a, b, out=out)
            ms = timer.median_ms(f, reps)
            rec({"op": name, "w": w, "in": "slab", "out": f"slab w={ob}"}, ms, 2 * in_b + out.words * 8, f,
                (am.KNOB_BINARY_WORDS, w, am.SLAB, am.SLAB))
            del out

        # u64 sum
        s = am.device_words(1)
        f = lambda: am.sum(a, ws=ws, out=s)
        ms = timer.median_ms(f, reps)
        rec({"op": "sum", "w": w, "in": "slab"}, ms, in_b + 8, f, (am.KNOB_SUM_BLOCKS, w, am.SLAB, 0))

        # bitstream variants at w = 11
        if w == 11:
            ab = am.convert(a, w, am.BITSTREAM)
            bb = am.convert(b, w, am.BITSTREAM)
            ib = ab.words * 8
            ms = timer.median_ms(lambda: am.filter(ab, am.LT, 1 << (w - 1), mask_out=mo), reps)
            rec({"op": "filter", "w": w, "in": "bitstream", "out": "mask"}, ms, ib + mask_words * 8)
            out = am.Column(n, w + 1, am.BITSTREAM)
            ms = timer.median_ms(lambda: am.binary(am.ADD, ab, bb, out=out), reps)
            rec({"op": "add", "w": w, "in": "bitstream", "out": "bitstream w=12"}, ms, 2 * ib + out.words * 8)
            f = lambda: am.sum(ab, ws=ws, out=s)
            ms = timer.median_ms(f, reps)
            rec({"op": "sum", "w": w, "in": "bitstream"}, ms, ib + 8, f, (am.KNOB_SUM_BLOCKS, w, am.BITSTREAM, 0))
            del ab, bb, out
        del a, b
        torch.cuda.empty_cache()
    for o in ops:
        o["ms"] = round(o["ms"], 4)
        o["GBs"] = round(o["GBs"], 1)
        o["pct_peak"] = round(100.0 * o["GBs"] / PEAK_GBS, 1)
    return ops


# The reference's result schema (src/main.cu:300, src/common_code.cuh:206-213):
# one line per timed launch, throughput = used_data_size / time in GiB/s,
# where used_data_size counts 8 bytes per element per input array
# ("uncompressed-equivalent"), plus the mask bytes for compaction. Files are
# split by workload as in the reference's data/ tree (one sweep file, one
# file per mask distribution and selectivity under comp/).
CSV_HEADER = "approach;element_count;bits;block_count;thread_count;time_ms;throughput;run"


def csv_place(o):
    """(file, approach, used_data_size) of a sweep record, or None."""
    n, op = o.get("n"), o["op"]
    if n is None:
        return None
    mask_b = (n + 7) // 8
    fmt = "nogaps" if o.get("in") == "bitstream" else "slabs"
    if op == "filter":
        return ("sweep", f"filter_{fmt}", 8 * n)
    if op == "add":
        return ("sweep", f"binary_op_{fmt}", 16 * n)
    if op == "and":
        return ("sweep", f"binary_op_{fmt}_and", 16 * n)
    if op == "sum":
        return ("sweep", f"hop_{fmt}", 8 * n)
    if "keys" in op:  # wide keys: the reference's sort-based group-by
        return ("sweep", "groupby_sort", 16 * n)
    if op.startswith("groupby") or op.endswith("groupby"):
        return ("sweep", "groupby_sm_array" if op.startswith("groupby") else "filter_groupby_sm_array", 16 * n)
    short = {"uniform": "uni", "clustered": "runs", "ref_1_cluster": "sc", "ref_8_clusters": "mc"}
    if op == "compressstore(mask)":
        return (f"comp/{short[o['mask']]}_{o['sel']}", "compressstore_slabs", 8 * n + mask_b)
    if op == "filter+compressstore":
        return (f"comp/pred_{o['sel']}", f"filter_compressstore_slabs_to_{o['out']}", 8 * n + mask_b)
    return None


def write_csv(ops, outdir):
    files = {}
    for o in ops:
        place = csv_place(o)
        if place is None or "times" not in o:
            continue
        f, approach, used = place
        lines = files.setdefault(f, [CSV_HEADER])
        for run, t in enumerate(o["times"]):
            thr = (used / (t / 1000.0)) / (1 << 30)
            lines.append(f"{approach};{o['n']};{o['w'] if 'w' in o else o.get('key_bits', 10)};-1;-1;{t:f};{thr:f};{run}")
    for f, lines in files.items():
        path = os.path.join(outdir, f)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
    return sorted(files)


def groupby_leg(am, torch, timer, reps):
    """256M rows: keys 10-bit Zipf(1.1) over 1024 ranks, values 20-bit uniform;
    exact SUM + COUNT, minimal-width key-sorted output; plus the chained
    filter (v > 2^19) variant."""
    n = 1 << 28
    p = 1.0 / np.arange(1, 1025) ** 1.1
    c = np.cumsum(p / p.sum())
    thr = np.minimum(np.floor(c * 4294967296.0), 4294967295).astype(np.uint64)
    thr[-1] = 4294967295
    keys = am.gen_from_cdf(42, thr.astype(np.uint32), n, 10, am.SLAB)
    vals = am.gen_uniform(43, n, 20, am.SLAB)
    ws = am.Workspace(0)
    out = []
    for label, pred in (("groupby sum+count", None), ("filter(v>2^19)->groupby", (am.GT, 1 << 19))):
        ms = timer.median_ms(lambda: am.groupby(keys, vals, value_pred=pred, ws=ws), reps)
        times = list(timer.last)
        _, _, _, res = am.groupby(keys, vals, value_pred=pred, ws=ws)
        torch.cuda.synchronize()
        r = am.u64(res)
        g = int(r[0])
        byt = keys.words * 8 + vals.words * 8 + g * 8 * 3
        out.append({"op": label, "rows": n, "n": n, "key_bits": 10, "groups": g, "sum_bits": int(r[1]),
                    "count_bits": int(r[2]), "ms": round(ms, 4), "GBs": round(gbytes(byt, ms), 1),
                    "pct_peak": round(100 * gbytes(byt, ms) / PEAK_GBS, 1), "times": times})
    del keys
    # wide keys (partitioned path): 24-bit uniform keys, 16-bit Zipf(1.1) ranks
    p16 = 1.0 / np.arange(1, 65537) ** 1.1
    c16 = np.cumsum(p16 / p16.sum())
    thr16 = np.minimum(np.floor(c16 * 4294967296.0), 4294967295).astype(np.uint64)
    thr16[-1] = 4294967295
    for label, kb, kcol in (("groupby 24-bit uniform keys", 24, lambda: am.gen_uniform(44, n, 24, am.SLAB)),
                            ("groupby 16-bit Zipf(1.1) keys", 16,
                             lambda: am.gen_from_cdf(45, thr16.astype(np.uint32), n, 16, am.SLAB))):
        keys = kcol()
        wsw = am.Workspace.for_groupby(n, kb, 20)
        ms = timer.median_ms(lambda: am.groupby(keys, vals, ws=wsw), reps)
        times = list(timer.last)
        _, _, _, res = am.groupby(keys, vals, ws=wsw)
        torch.cuda.synchronize()
        r = am.u64(res)
        g = int(r[0])
        outb = (am.packed_words(am.SLAB, kb, g) + am.packed_words(am.SLAB, int(r[1]), g)
                + am.packed_words(am.SLAB, int(r[2]), g)) * 8
        byt = keys.words * 8 + vals.words * 8 + outb
        out.append({"op": label, "rows": n, "n": n, "key_bits": kb, "groups": g, "sum_bits": int(r[1]),
                    "count_bits": int(r[2]), "ms": round(ms, 4), "GBs": round(gbytes(byt, ms), 1),
                    "pct_peak": round(100 * gbytes(byt, ms) / PEAK_GBS, 1), "times": times})
        del keys, wsw
        torch.cuda.empty_cache()
    return out


def matmul_leg(am, torch, timer, reps):
    """A, B 8192x8192 7-bit (slab packed), exact i32 via int8 MFMA, C re-packed
    at 27 bits; plus 256 uneven matrices (M, N, K ~ U[64, 2048]) in one batch."""
    M = N = K = 8192
    bits = 7
    words = am.matrix_words(M, K, bits)
    A = am.gen_uniform(5, words * 9, 7, am.SLAB)  # any 7-bit slab words of the right count
    B = am.gen_uniform(6, words * 9, 7, am.SLAB)
    prob = [(A.buf, B.buf, M, N, K)]
    plan = am.MatmulPlan(bits, prob)  # header uploaded once; each step is the three launches
    outs = plan.outs
    plan.run()
    torch.cuda.synchronize()
    ms = timer.median_ms(plan.run, max(2, reps // 3), warm=1)
    tops = 2.0 * M * N * K / (ms * 1e-3) / 1e12
    res = [{"op": "matmul u7 8192^3 (unpack+gemm+pack)", "ms": round(ms, 3), "TOPS": round(tops, 1),
            "pct_int8_peak": round(100 * tops / 5000.0, 1), "out_bits": outs[0][2]}]
    rng = np.random.default_rng(7)
    probs = []
    flops = 0
    for i in range(256):
        m, nn, k = (int(x) for x in rng.integers(64, 2049, 3))
        wa, wb = am.matrix_words(m, k, bits), am.matrix_words(k, nn, bits)
        a = am.gen_uniform(100 + i, wa * 9, 7, am.SLAB)
        b = am.gen_uniform(1000 + i, wb * 9, 7, am.SLAB)
        probs.append((a.buf, b.buf, m, nn, k))
        flops += 2.0 * m * nn * k
    bplan = am.MatmulPlan(bits, probs)
    bplan.run()
    torch.cuda.synchronize()
    ms = timer.median_ms(bplan.run, max(2, reps // 3), warm=1)
    tops = flops / (ms * 1e-3) / 1e12
    res.append({"op": "matmul u7 batch of 256 uneven", "ms": round(ms, 3), "TOPS": round(tops, 1),
                "pct_int8_peak": round(100 * tops / 5000.0, 1)})
    return res


def libbase_leg(am, torch, timer, reps):
    """The reference's GPU library baselines, run on this GPU: hipCUB (rocPRIM)
    over one u64 per element -- DeviceSelect::Flagged with the bitmask
    (src/benchs/bench_compressstore_baseline_cub.cu:29-35) and DeviceReduce::Sum
    (src/benchs/bench_hop_baseline_cub.cu:32-38) -- beside the library's
    operators on the same 2^28 values (uniform 11-bit, the headline's v < 1024
    selection): unpacked (u64 in and out) and packed (11-bit slabs)."""
    from amethyst_amd import baseline as hb
    n = N_FULL
    col64 = am.gen_uniform(42, n, 11, am.BASELINE)
    slab = am.gen_uniform(42, n, 11, am.SLAB)
    mask = am.filter(slab, am.LT, 1024)
    sel = hb.SelectFlagged(col64.buf, mask, n)
    flags8 = (col64.buf < 1024).to(torch.uint8)  # the same selection, one byte per element
    sel8 = hb.SelectFlagged(col64.buf, flags8, n, bytes=True)
    red = hb.ReduceSum(col64.buf, n)
    ws = am.Workspace(n)
    stats = am.device_words(2)
    tot = am.device_words(1)
    out64 = am.Column(n, 64, am.BASELINE)
    outs = am.Column(n, 11, am.SLAB)
    sel.run()
    sel8.run()
    red.run()
    am.compressstore(col64, 64, am.BASELINE, mask=mask, ws=ws, out=out64, stats=stats)
    am.sum(col64, ws=ws, out=tot)
    torch.cuda.synchronize()
    k = int(am.u64(sel.count)[0])
    same = (k == int(am.u64(stats)[0]) == int(am.u64(sel8.count)[0])
            and int(am.u64(red.out)[0]) == int(am.u64(tot)[0])
            and bool(torch.equal(sel.out[:k], out64.buf[:k])) and bool(torch.equal(sel8.out[:k], out64.buf[:k])))
    mw = (n + 63) // 64 * 8
    rows = []

    def rec(op, fn, byt):
        ms = timer.median_ms(fn, reps)
        rows.append({"op": op, "n": n, "selected": k, "ms": round(ms, 4), "GBs": round(gbytes(byt, ms), 1),
                     "pct_peak": round(100 * gbytes(byt, ms) / PEAK_GBS, 1)})

    rec("hipcub DeviceSelect::Flagged, u64 + bitmask -> u64", sel.run, n * 8 + mw + k * 8)
    rec("hipcub DeviceSelect::Flagged, u64 + flag bytes -> u64", sel8.run, n * 8 + n + k * 8)
    rec("compressstore(mask), u64 -> u64", lambda: am.compressstore(col64, 64, am.BASELINE, mask=mask, ws=ws,
                                                                    out=out64, stats=stats), n * 8 + mw + k * 8)
    rec("compressstore(mask), 11-bit slab -> slab", lambda: am.compressstore(slab, 11, am.SLAB, mask=mask, ws=ws,
                                                                             out=outs, stats=stats),
        slab.words * 8 + mw + am.packed_words(am.SLAB, 11, k) * 8)
    rec("hipcub DeviceReduce::Sum, u64", red.run, n * 8)
    rec("sum, u64", lambda: am.sum(col64, ws=ws, out=tot), n * 8)
    rec("sum, 11-bit slab", lambda: am.sum(slab, ws=ws, out=tot), slab.words * 8)
    # time against hipCUB's faster form of the same operation
    bsel = min(rows[0]["ms"], rows[1]["ms"])
    for r in rows:
        base = bsel if "Select" in r["op"] or "compressstore" in r["op"] else rows[4]["ms"]
        r["speedup_vs_hipcub"] = round(base / r["ms"], 2)
    del col64, slab, out64, outs, sel, sel8, flags8, red
    torch.cuda.empty_cache()
    return {"results_match": same, "rows": rows}


def batched_leg(am, torch, reps):
    """512 columns, lengths log-uniform in [4K, 64M], widths U[3,31], values
    uniform / Zipf(1.2) / sorted+noise; per column filter -> compressstore ->
    add second column -> sum. Plain: one operator launch at a time per column
    (count read back between stages). Batched: planner (HIP features + MLP, or
    the analytic plan without trained weights) then one launch for the whole
    batch (each column's tile size per its plan), also replayed from a
    captured graph."""
    from amethyst_amd import planner
    cols = []
    for i in range(512):
        spec = planner.column_spec(50_000 + i)
        cols.append(planner.make_inputs(50_000 + i, spec))
    torch.cuda.synchronize()
    rows = sum(c[0].n for c in cols)
    # plain plan
    planner.run_plain(cols)
    t0 = time.perf_counter()
    plain_totals = planner.run_plain(cols)
    t_plain = time.perf_counter() - t0
    # planning
    t0 = time.perf_counter()
    feat = planner.features(cols)
    try:
        pl = planner.Planner()
        plans, _ = pl.choose(feat)
        kind = "learned"
    except (OSError, FileNotFoundError):
        plans = np.full(len(cols), planner.ANALYTIC, dtype=np.int32)
        kind = "analytic (no trained weights)"
    t_plan = time.perf_counter() - t0
    fr = feat.cpu().numpy()
    outs, stats, t_batch, redo = planner.run_batch(cols, list(plans), fr, reps=max(3, reps // 2))
    ok = [int(s[2]) for s in stats] == plain_totals
    # algorithmic bytes of the fused chain: both inputs read once, the
    # compacted sums written at each column's planned format and width
    byt = 0
    for (a, b, _), o, st in zip(cols, outs, stats):
        byt += a.words * 8 + b.words * 8 + am.packed_words(o.fmt, o.bits, int(st[0])) * 8
    # graph replay of the batched launch (same plans; overflowed columns at w+1)
    spec, variants = [], []
    for i, p in enumerate(plans):
        a, b, pred = cols[i]
        fmt, wr, var = planner.PLANS[p]
        w = planner.plan_width(wr, a.bits, planner.est_bits(fr[i]))
        if int(stats[i][1]):
            w = a.bits + 1
        spec.append((a, b, pred, fmt, w))
        variants.append(var)
    bt = am.PipelineBatch(spec, variant=variants)
    g = bt.capture()
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nrep = max(3, reps // 2)
    for _ in range(nrep):
        g.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / nrep
    res_ok = [int(x[2]) for x in bt.results()[1]] == plain_totals
    return {"columns": len(cols), "rows": rows, "plain_s": round(t_plain, 5), "batched_s": round(t_batch, 5),
            "graph_s": round(t_graph, 5), "plan_s": round(t_plan, 5), "planner": kind,
            "bytes": byt, "GBs": round(byt / t_batch / 1e9, 1), "pct_peak": round(100 * byt / t_batch / 1e9 / PEAK_GBS, 1),
            "speedup_batched": round(t_plain / t_batch, 2), "speedup_graph": round(t_plain / t_graph, 2),
            "speedup_incl_planning": round(t_plain / (t_batch + t_plan), 2),
            "fallback_reruns": redo, "totals_match_plain": bool(ok and res_ok)}


def run_case(am, torch, name, steps):
    """One workload, `steps` launches, for profiling: the headline compaction,
    fused filter + compressstore into same-width slabs at 50 % selectivity
    (w = 4, 11), the 512-column batched chain (one launch), the wide-key
    group-by (24-bit uniform / 16-bit Zipf keys), ADD of two w=4 slabs into a
    w=5 slab or of two 11-bit bitstreams into a 12-bit bitstream, the sum of
    an 11-bit bitstream, or a plain 1 GiB device copy (the reference point of
    the HBM byte counters)."""
    n = N_FULL
    if name in ("add_slab_w4", "add_bitstream_w11"):
        w, fmt = (4, am.SLAB) if name == "add_slab_w4" else (11, am.BITSTREAM)
        a, b = am.gen_uniform(11, n, w, fmt), am.gen_uniform(12, n, w, fmt)
        out = am.Column(n, w + 1, fmt)
        fn = lambda: am.binary(am.ADD, a, b, out=out)
        byt = (a.words + b.words + out.words) * 8
    elif name == "sum_bitstream_w11":
        a = am.gen_uniform(21, n, 11, am.BITSTREAM)
        ws, s = am.Workspace(0), am.device_words(1)
        fn = lambda: am.sum(a, ws=ws, out=s)
        byt = a.words * 8 + 8
    elif name == "copy_1gib":
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        fn, byt = (lambda: dst.copy_(src)), 2 << 30
    elif name == "pipeline":
        from amethyst_amd import planner
        cols = [planner.make_inputs(50_000 + i, planner.column_spec(50_000 + i)) for i in range(512)]
        feat = planner.features(cols)
        plans, _ = planner.Planner().choose(feat)
        fr = feat.cpu().numpy()
        spec, var = [], []
        for i, p in enumerate(plans):
            fmt, wr, v = planner.PLANS[p]
            spec.append((cols[i][0], cols[i][1], cols[i][2], fmt, planner.plan_width(wr, cols[i][0].bits,
                                                                                      planner.est_bits(fr[i]))))
            var.append(v)
        bt = am.PipelineBatch(spec, variant=var)
        bt.run()
        torch.cuda.synchronize()
        st = bt.results()[1]
        byt = sum(a.words * 8 + b.words * 8 + am.packed_words(o.fmt, o.bits, int(x[0])) * 8
                  for (a, b, _), o, x in zip(cols, bt.outs, st))
        fn = bt.run
    elif name in ("groupby_u24", "groupby_z16"):
        # wide-key group-by (partitioned path): 2^28 rows, 20-bit uniform
        # values, 24-bit uniform keys or 16-bit Zipf(1.1) ranks
        vals = am.gen_uniform(43, n, 20, am.SLAB)
        if name == "groupby_u24":
            kb, keys = 24, am.gen_uniform(44, n, 24, am.SLAB)
        else:
            p16 = 1.0 / np.arange(1, 65537) ** 1.1
            c16 = np.cumsum(p16 / p16.sum())
            thr16 = np.minimum(np.floor(c16 * 4294967296.0), 4294967295).astype(np.uint64)
            thr16[-1] = 4294967295
            kb, keys = 16, am.gen_from_cdf(45, thr16.astype(np.uint32), n, 16, am.SLAB)
        wsw = am.Workspace.for_groupby(n, kb, 20)
        fn = lambda: am.groupby(keys, vals, ws=wsw)
        _, _, _, res = fn()
        torch.cuda.synchronize()
        r = am.u64(res)
        g = int(r[0])
        byt = keys.words * 8 + vals.words * 8 + (am.packed_words(am.SLAB, kb, g) + am.packed_words(
            am.SLAB, int(r[1]), g) + am.packed_words(am.SLAB, int(r[2]), g)) * 8
    else:
        w = {"headline": 11, "compress_slab_w4": 4, "compress_slab_w11": 11}[name]
        col = am.gen_uniform(42, n, w, am.SLAB)
        ws = am.Workspace(n)
        stats = am.device_words(2)
        if name == "headline":
            out = am.Column(n, 32, am.U32)
            mo = am.device_words((n + 63) // 64)
            fn = lambda: am.compressstore(col, 32, am.U32, pred=(am.LT, 1024), ws=ws, out=out, stats=stats, mask_out=mo)
        else:
            out = am.Column(n, w, am.SLAB)
            fn = lambda: am.compressstore(col, w, am.SLAB, pred=(am.LT, 1 << (w - 1)), ws=ws, out=out, stats=stats)
        fn()
        torch.cuda.synchronize()
        k = int(am.u64(stats)[0])
        byt = col.words * 8 + out.words * 0 + am.packed_words(out.fmt, out.bits, k) * 8
        if name == "headline":
            byt += (n + 63) // 64 * 8
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    print(json.dumps({"case": name, "launches": steps, "algorithmic_bytes_per_launch": byt}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the operators run on a single GPU)", file=sys.stderr)
        sys.exit(2)
    import torch
    import amethyst_amd as am
    am.lib()
    torch.cuda.init()
    timer = Timer(torch)
    if args.case:
        run_case(am, torch, args.case, args.steps)
        return
    if args.profile_copy:
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- headline workload
    n, w, c = N_FULL, 11, 1024
    col = am.gen_uniform(42, n, w, am.SLAB)
    out = am.Column(n, 32, am.U32)
    ws = am.Workspace(n)
    stats = am.device_words(2)
    mask_out = am.device_words((n + 63) // 64)

    def step():
        am.compressstore(col, 32, am.U32, pred=(am.LT, c), ws=ws, out=out, stats=stats, mask_out=mask_out)

    print(f"headline: compressstore n={n} w={w}", file=sys.stderr, flush=True)
    step()
    torch.cuda.synchronize()
    ws.check()
    k = int(am.u64(stats)[0])
    byt = col.words * 8 + 4 * k + ((n + 63) // 64) * 8
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-launch kernel time with HIP events on the launch stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a_, b_ in evs:
        a_.record()
        step()
        b_.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ws.check()
    kms = [a_.elapsed_time(b_) for a_, b_ in evs]
    ms_step = wall * 1e3 / args.steps
    avg_kernel = sum(kms) / len(kms)
    value = gbytes(byt, ms_step)
    achieved = gbytes(byt, avg_kernel)
    traffic = read_traffic("k_compress_w11_u32_traffic.txt")

    legs = set() if args.quick else set(args.legs.split(","))
    ops = sweep(am, torch, timer, args.reps) if "sweep" in legs else []
    extra = {}
    if "groupby" in legs:
        print("group-by", file=sys.stderr, flush=True)
        ops += groupby_leg(am, torch, timer, args.reps)
    if "matmul" in legs:
        print("matmul", file=sys.stderr, flush=True)
        ops += matmul_leg(am, torch, timer, args.reps)
    if "batched" in legs:
        print("batched", file=sys.stderr, flush=True)
        extra["batched"] = batched_leg(am, torch, args.reps)
    if "libbase" in legs:
        print("library baseline (hipCUB)", file=sys.stderr, flush=True)
        extra["library_baseline"] = libbase_leg(am, torch, timer, args.reps)
    # measured copy rates on this box, 1 GiB device to device: the library's
    # streaming copy (16-byte non-temporal loads and stores) and torch's copy_
    src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    copy_ms = timer.median_ms(lambda: am.stream_copy(src, dst), max(5, args.reps))
    copy_gbs = gbytes(2 * src.numel() * 8, copy_ms)
    tcopy_ms = timer.median_ms(lambda: dst.copy_(src), max(5, args.reps))
    tcopy_gbs = gbytes(2 * src.numel() * 8, tcopy_ms)
    del src, dst
    print("cpu baseline", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(am, torch, col, c)

    if args.csv:
        extra["csv_files"] = write_csv(ops, args.csv)
    for o in ops:
        o.pop("times", None)

    line = {
        "metric": "HBM GB/s (% of peak) per op; batched+learned plan vs plain plan speedup",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (PCG32 seed 42, generated on the GPU)",
        "config": {"workload": "filter(v<1024)+compressstore, N=2^28 uniform 11-bit values, slab in, "
                               "u32 out + bitmask, one MI355X", "n": n, "bits": w, "selected": k,
                   "bytes_per_step": byt},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_compress_direct<PRED=1,OVF=0,VB=2,WIDE=0>", "avg_kernel_ms": round(avg_kernel, 5)},
        "copy_peak": {"value": round(copy_gbs, 1), "unit": "GB/s",
                      "kernel": "am_stream_copy of 1 GiB (16-byte non-temporal loads and stores)",
                      "headline_frac_of_copy": round(achieved / copy_gbs, 4),
                      "torch_copy": round(tcopy_gbs, 1)},
        "cpu_baseline": cpu,
        "ops": ops,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
