"""Autotuner for the launch-shape table (am_tune_* in include/amethyst.h).

For every (knob, width, input format, output format) it times the operator
at the tuning size with the built-in launch shape and with each candidate
value of the knob (HIP events on the launch stream, median of `reps`
launches), and keeps a candidate only where it beats the built-in shape by
more than `min_gain`. The result is a text table the library reads
(amethyst_amd/tune_table.txt, loaded by amethyst_amd.lib()). The reference
picks its launch shape per width by sweeping grid x block sizes at benchmark
time (src/common_code.cuh:179-225, src/utils.cuh:209-257); here the sweep
runs once per device and its choice is stored.

    python -m amethyst_amd.autotune [--n 268435456] [--reps 10] [--out amethyst_amd/tune_table.txt]

Knobs and candidates:
  filter words per lane        slab / u32 inputs, predicate v < 2^(w-1)
  compaction chunks per wave   slab -> u32 and slab -> slab (same width), v < 2^(w-1)
  binary words per thread      slab add (w -> w+1 bits), applies to every operator
  sum blocks                   slab / u32 / bitstream inputs
  group-by blocks              slab keys of 4..12 bits, 20-bit slab values
"""
import argparse
import statistics
import sys
import time


import amethyst_amd as am

CANDIDATES = {
    am.KNOB_FILTER_WORDS: (1, 2, 4, 8),
    am.KNOB_COMPRESS_CHUNKS: (1, 2, 3, 4, 5, 6, 7, 8),
    am.KNOB_BINARY_WORDS: (2, 4, 8),
    am.KNOB_SUM_BLOCKS: (256, 512, 1024, 2048, 4096, 8192),
    am.KNOB_GROUPBY_BLOCKS: (128, 256, 512, 1024, 2048),
}
NAMES = {am.KNOB_FILTER_WORDS: "filter words", am.KNOB_COMPRESS_CHUNKS: "compress chunks",
         am.KNOB_BINARY_WORDS: "binary words", am.KNOB_SUM_BLOCKS: "sum blocks",
         am.KNOB_GROUPBY_BLOCKS: "group-by blocks", am.KNOB_PIPELINE_TILE: "pipeline tile"}


def median_ms(torch, fn, reps, warm=2):
    for _ in range(warm):
        fn()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return statistics.median(a.elapsed_time(b) for a, b in evs)


def configs(torch, n, widths):
    """Yields (knob, bits, in_fmt, out_fmt, fn) with fn one launch of the operator."""
    ws = am.Workspace(n)
    stats = am.device_words(2)
    mask = am.device_words((n + 63) // 64)
    s = am.device_words(1)
    for w in widths:
        print(f"autotune w={w}", file=sys.stderr, flush=True)
        fmts = [am.SLAB] + ([am.U32] if w <= 32 else [])
        for fi in fmts:
            a = am.gen_uniform(42, n, w, fi)
            c = 1 << (w - 1)
            yield am.KNOB_FILTER_WORDS, w, fi, 0, (lambda: am.filter(a, am.LT, c, mask_out=mask))
            yield am.KNOB_SUM_BLOCKS, w, fi, 0, (lambda: am.sum(a, ws=ws, out=s))
            if fi == am.SLAB:
                for fo, ob in ((am.U32, min(w, 32)), (am.SLAB, w)):
                    out = am.Column(n, ob, fo)
                    yield am.KNOB_COMPRESS_CHUNKS, w, fi, fo, (
                        lambda: am.compressstore(a, ob, fo, pred=(am.LT, c), ws=ws, out=out, stats=stats))
                    del out
                if w < 64:
                    b = am.gen_uniform(43, n, w, fi)
                    out = am.Column(n, w + 1, am.SLAB)
                    yield am.KNOB_BINARY_WORDS, w, fi, am.SLAB, (lambda: am.binary(am.ADD, a, b, out=out))
                    del b, out
            del a
            torch.cuda.empty_cache()
        if w >= 3 and 64 % w != 0 and w <= 32:  # bitstreams whose width divides 64 are slabs
            a = am.gen_uniform(42, n, w, am.BITSTREAM)
            yield am.KNOB_SUM_BLOCKS, w, am.BITSTREAM, 0, (lambda: am.sum(a, ws=ws, out=s))
            del a
    vals = am.gen_uniform(43, n, 20, am.SLAB)
    gws = am.Workspace(0)
    for kb in (4, 8, 10, 12):
        keys = am.gen_uniform(44, n, kb, am.SLAB)
        yield am.KNOB_GROUPBY_BLOCKS, kb, am.SLAB, am.SLAB, (lambda: am.groupby(keys, vals, ws=gws))
        del keys


def tune(n=1 << 28, reps=10, min_gain=0.02, widths=None, out=None, log=None):
    import torch
    am.lib()
    am.tune_clear()
    widths = widths or list(range(1, 33)) + [40, 48, 64]
    rows = []
    t_start = time.time()
    for knob, bits, fi, fo, fn in configs(torch, n, widths):
        am.tune_set(knob, bits, fi, fo, 0)
        t0 = median_ms(torch, fn, reps)
        best, tb = 0, t0
        times = {}
        for v in CANDIDATES[knob]:
            am.tune_set(knob, bits, fi, fo, v)
            t = median_ms(torch, fn, reps)
            times[v] = t
            if t < tb:
                best, tb = v, t
        keep = best and tb < t0 * (1.0 - min_gain)
        if keep:  # confirm: the built-in shape and the candidate timed again, alternately
            d, c = [], []
            for _ in range(3):
                am.tune_set(knob, bits, fi, fo, 0)
                d.append(median_ms(torch, fn, reps))
                am.tune_set(knob, bits, fi, fo, best)
                c.append(median_ms(torch, fn, reps))
            t0, tb = statistics.median(d), statistics.median(c)
            keep = tb < t0 * (1.0 - min_gain)
        am.tune_set(knob, bits, fi, fo, best if keep else 0)
        rows.append((knob, bits, fi, fo, best if keep else 0, t0, tb if keep else t0, times))
        if log:
            log(f"{NAMES[knob]:16s} w={bits:2d} in={fi} out={fo}: default {t0:.4f} ms, best {best} {tb:.4f} ms"
                f"{'  kept' if keep else ''}")
    if out:
        write_table(rows, out, n, reps, time.time() - t_start)
    return rows


def write_table(rows, path, n, reps, secs):
    import torch
    dev = torch.cuda.get_device_name(0)
    kept = [r for r in rows if r[4]]
    with open(path, "w") as f:
        f.write("# Launch-shape table of the packed-column operators (am_tune_load):\n")
        f.write("#   knob bits in_format out_format value\n")
        f.write("# knobs: 0 filter words per lane, 1 compaction chunks per wave, 2 binary words per thread,\n")
        f.write("#        3 sum blocks, 4 group-by blocks; formats: 0 baseline 1 slab 2 bitstream 3 tight 4 u32\n")
        f.write(f"# written by amethyst_amd/autotune.py on {dev}: n = {n}, median of {reps} launches per point,\n")
        f.write(f"# {len(rows)} points swept in {secs:.0f} s, {len(kept)} kept (> 2 % faster than the built-in shape\n")
        f.write("# in the sweep and again in a confirming re-measurement, alternating with the built-in shape)\n")
        for knob, bits, fi, fo, v, t0, tb, _ in rows:
            if v:
                f.write(f"{knob} {bits} {fi} {fo} {v}   # {NAMES[knob]}: {t0:.4f} -> {tb:.4f} ms\n")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=1 << 28)
    p.add_argument("--reps", type=int, default=10)
    p.add_argument("--min-gain", type=float, default=0.02)
    p.add_argument("--widths", default=None, help="comma-separated widths (default 1..32, 40, 48, 64)")
    p.add_argument("--out", default=am.TUNE_TABLE)
    args = p.parse_args()
    widths = [int(x) for x in args.widths.split(",")] if args.widths else None
    rows = tune(args.n, args.reps, args.min_gain, widths, args.out, log=lambda m: print(m, flush=True))
    kept = sum(1 for r in rows if r[4])
    print(f"{kept} of {len(rows)} points tuned -> {args.out}")


if __name__ == "__main__":
    main()
