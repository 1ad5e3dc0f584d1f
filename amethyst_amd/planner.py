"""Learned planner for batched operator chains (filter -> compressstore(a+b) -> sum).

Plan space, per column (none of it changes any integer arithmetic):
  * output format of s: slab or bitstream;
  * output width: w+1 (the add rule, always safe), or the sampled maximum
    width "est" / "est+1" (narrower, fewer bytes; a value that does not fit
    is detected on the device and the column re-runs at w+1);
  * kernel variant: 8192- or 4096-element tiles (J = 32 or 16), per column
    within the batch's one launch.

The planner is an MLP (32 -> 64 -> 64 -> 12) over per-column statistics that
the HIP kernel k_plan_stats computes; it predicts each plan's log-time
relative to the best plan and runs as the hand-written HIP kernel
k_plan_mlp. When the predicted best beats the runner-up by less than a
margin, the column takes the analytic plan (bitstream, width w+1, J = 32:
fewest bytes among the always-safe plans). Training data are the timings of
every plan on seeded columns, produced by this module's own sweep
(`collect`); `evaluate` measures regret against the exhaustively best plan
on held-out seeds.

    python -m amethyst_amd.planner collect  --n 3000 --seed 1000 --out plan_train.npz
    python -m amethyst_amd.planner train    --data plan_train.npz --out amethyst_amd/planner_weights.npz
    python -m amethyst_amd.planner evaluate --n 10000 --seed 900000 --out plan_eval.json
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import (BITSTREAM, LT, SLAB, Column, PipelineBatch, _check, _stream, device_words, lib, u64)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "planner_weights.npz")

PLANS = [(fmt, wr, var) for var in (32, 16) for fmt in (SLAB, BITSTREAM) for wr in ("w+1", "est", "est+1")]
NPLANS = len(PLANS)
ANALYTIC = PLANS.index((BITSTREAM, "w+1", 32))
NFEAT, NHID = 32, 64


def plan_width(rule, w, est):
    if rule == "w+1":
        return w + 1
    return max(1, min(w + 1, est + (1 if rule == "est+1" else 0)))


# ----------------------------------------------------------------- columns

def zipf_thresholds(ranks, s=1.2):
    p = 1.0 / np.arange(1, ranks + 1) ** s
    c = np.cumsum(p / p.sum())
    thr = np.minimum(np.floor(c * 4294967296.0), 4294967295).astype(np.uint64)
    thr[-1] = 4294967295
    return thr.astype(np.uint32)


def gen_column(seed, n, w, dist, fmt=SLAB):
    """Seeded GPU-generated column: 0 uniform, 1 Zipf(1.2) ranks, 2 sorted + noise."""
    import amethyst_amd as am
    if dist == 0:
        return am.gen_uniform(seed, n, w, fmt)
    if dist == 1:
        return am.gen_from_cdf(seed, zipf_thresholds(min(1 << w, 65536)), n, w, fmt)
    col = Column(n, w, fmt)
    c = col.c()
    _check(lib().am_gen_sorted_noise(seed, max(1, w - 6), ctypes.byref(c), _stream()), "am_gen_sorted_noise")
    return col


def column_spec(seed, nmin=4096, nmax=1 << 26):
    """Length log-uniform in [nmin, nmax], width uniform in [3, 31], dist by seed."""
    rng = np.random.default_rng(seed)
    n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    w = int(rng.integers(3, 32))
    dist = seed % 3
    q = float(rng.uniform(0.02, 0.98))
    c = max(1, int(q * (1 << w)))
    return n, w, dist, c


def make_inputs(seed, spec):
    n, w, dist, c = spec
    a = gen_column(2 * seed, n, w, dist)
    b = gen_column(2 * seed + 1, n, w, 0)
    return a, b, (LT, c)


# ----------------------------------------------------------------- features

def features(cols):
    """cols: list of (a, b, pred) -> device float tensor [len][32] (HIP k_plan_stats)."""
    import amethyst_amd as am
    torch = am._torch()
    n = len(cols)
    arr = (am.am_pipeline_col * n)()
    for i, (a, b, pred) in enumerate(cols):
        arr[i] = am.am_pipeline_col(a.c(), b.c(), a.c(), am.am_predicate(pred[0], 0, pred[1]), None)
    feat = torch.zeros(n * NFEAT, dtype=torch.float32, device="cuda")
    scratch = device_words((int(lib().am_plan_scratch_bytes(n)) + 7) // 8)
    _check(lib().am_plan_features(arr, n, ctypes.c_void_p(feat.data_ptr()), ctypes.c_void_p(scratch.data_ptr()),
                                  _stream()), "am_plan_features")
    torch.cuda.synchronize()
    return feat.view(n, NFEAT)


def est_bits(feat_row):
    return max(1, int(round(float(feat_row[19]) * 33.0)))


# ----------------------------------------------------------------- execution

def run_batch(cols, plans, feats, reps=0):
    """Run columns with per-column plans in one launch (each column's tile
    size is part of its plan); columns whose values did not fit the planned
    width re-run at w+1. Returns (outs, stats, seconds per timed rep, reruns)."""
    import amethyst_amd as am
    torch = am._torch()
    spec, variants = [], []
    for i, p in enumerate(plans):
        a, b, pred = cols[i]
        fmt, wr, var = PLANS[p]
        spec.append((a, b, pred, fmt, plan_width(wr, a.bits, est_bits(feats[i]))))
        variants.append(var)
    bt = PipelineBatch(spec, variant=variants)
    bt.run()
    torch.cuda.synchronize()
    secs = None
    if reps:
        t0 = time.perf_counter()
        for _ in range(reps):
            bt.run()
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / reps
    o, st = bt.results()
    outs, stats = list(o), [st[i] for i in range(len(cols))]
    redo = [i for i in range(len(cols)) if int(stats[i][1])]
    if redo:  # verified fallback: values wider than the chosen width
        spec = [(cols[i][0], cols[i][1], cols[i][2], PLANS[plans[i]][0], cols[i][0].bits + 1) for i in redo]
        bt2 = PipelineBatch(spec, variant=32)
        bt2.run()
        torch.cuda.synchronize()
        o2, st2 = bt2.results()
        for j, i in enumerate(redo):
            outs[i], stats[i] = o2[j], st2[j]
    return outs, stats, secs, len(redo)


def time_plans(a, b, pred, feat_row, reps=3):
    """Median kernel time (s) of every plan on one column, including the
    fallback re-run when a narrow width overflows."""
    import amethyst_amd as am
    torch = am._torch()
    est = est_bits(feat_row)
    times = np.zeros(NPLANS)
    safe = {}
    for p, (fmt, wr, var) in enumerate(PLANS):
        bt = PipelineBatch([(a, b, pred, fmt, plan_width(wr, a.bits, est))], variant=var)
        bt.run()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for s_, e_ in ev:
            s_.record()
            bt.run()
            e_.record()
        torch.cuda.synchronize()
        t = float(np.median([s_.elapsed_time(e_) for s_, e_ in ev])) * 1e-3
        ovf = int(bt.results()[1][0][1])
        times[p] = t
        if wr == "w+1":
            safe[(fmt, var)] = t
        if ovf:
            times[p] = t + np.nan  # filled below once the safe plan is timed
        del bt
    for p, (fmt, wr, var) in enumerate(PLANS):
        if np.isnan(times[p]):
            # re-time this plan's own cost and add the safe re-run
            bt = PipelineBatch([(a, b, pred, fmt, plan_width(wr, a.bits, est))], variant=var)
            bt.run()
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
            bt.run()
            ev[1].record()
            torch.cuda.synchronize()
            times[p] = ev[0].elapsed_time(ev[1]) * 1e-3 + safe[(fmt, 16)]
    return times


def run_plain(cols):
    """The reference-style chain, one operator launch at a time per column:
    filter -> compressstore(a) -> compressstore(b) -> add -> sum, reading the
    selected count back to the host between stages (as the reference's
    compressstore bench does, src/benchs/bench_compressstore_slabs.cu:59-62)."""
    import amethyst_amd as am
    torch = am._torch()
    totals = []
    ws = am.Workspace(max(c[0].n for c in cols))
    stats = am.device_words(2)
    for a, b, pred in cols:
        m = am.filter(a, pred[0], pred[1])
        ca, _, _ = am.compressstore(a, a.bits, SLAB, mask=m, ws=ws, stats=stats)
        cb, _, _ = am.compressstore(b, b.bits, SLAB, mask=m, ws=ws, stats=stats)
        k = int(u64(stats)[0])
        ca.n = cb.n = k
        s = am.binary(am.ADD, ca, cb, out_bits=a.bits + 1, out_fmt=SLAB) if k else None
        tot = am.sum(s, ws=ws) if k else None
        totals.append(tot)
    torch.cuda.synchronize()
    return [int(u64(t)[0]) if t is not None else 0 for t in totals]


# ----------------------------------------------------------------- MLP

def _init(rng, fan_in, fan_out):
    return (rng.standard_normal((fan_out, fan_in)) * math.sqrt(2.0 / fan_in)).astype(np.float32)


def train(X, T, epochs=400, lr=3e-3, seed=0):
    """X: [N,32] features, T: [N,P] plan times -> weights dict (numpy).

    Targets are log(t_p / t_best); loss is MSE, with extra weight on the plans
    near the best so the argmin is what the fit gets right."""
    import torch
    torch.manual_seed(seed)
    X = np.asarray(X, np.float32)
    T = np.asarray(T, np.float64)
    Y = np.log(T / T.min(axis=1, keepdims=True)).astype(np.float32)
    mean = X.mean(axis=0)
    std = X.std(axis=0) + 1e-6
    Xn = torch.tensor((X - mean) / std)
    Yt = torch.tensor(Y)
    W = torch.exp(-4.0 * Yt)  # emphasise near-best plans
    model = torch.nn.Sequential(torch.nn.Linear(NFEAT, NHID), torch.nn.ReLU(), torch.nn.Linear(NHID, NHID),
                                torch.nn.ReLU(), torch.nn.Linear(NHID, Y.shape[1]))
    opt = torch.optim.Adam(model.parameters(), lr=lr, weight_decay=1e-5)
    n = len(Xn)
    for ep in range(epochs):
        perm = torch.randperm(n)
        for i in range(0, n, 256):
            b = perm[i:i + 256]
            pred = model(Xn[b])
            loss = ((pred - Yt[b]) ** 2 * (0.25 + W[b])).mean()
            opt.zero_grad()
            loss.backward()
            opt.step()
    sd = [p.detach().numpy().astype(np.float32) for p in model.parameters()]
    flat = np.concatenate([sd[0].ravel(), sd[1], sd[2].ravel(), sd[3], sd[4].ravel(), sd[5]])
    return {"weights": flat, "mean": mean.astype(np.float32), "std": std.astype(np.float32),
            "nplans": np.int32(Y.shape[1])}


class Planner:
    """Loads trained weights and picks plans with the HIP MLP kernel."""

    def __init__(self, path=WEIGHTS, min_margin=None):
        import amethyst_amd as am
        torch = am._torch()
        z = np.load(path)
        if min_margin is None:
            min_margin = float(z["margin"]) if "margin" in z.files else 0.02
        self.P = int(z["nplans"])
        self.w = torch.tensor(z["weights"], device="cuda")
        self.mean = torch.tensor(z["mean"], device="cuda")
        self.std = torch.tensor(z["std"], device="cuda")
        self.min_margin = float(min_margin)

    def choose(self, feat, scores=False):
        """feat: device [n][32] -> (plans numpy int32, scores or None)."""
        import amethyst_amd as am
        torch = am._torch()
        n = feat.shape[0]
        fb = torch.full((n,), ANALYTIC, dtype=torch.int32, device="cuda")
        plan = torch.zeros(n, dtype=torch.int32, device="cuda")
        sc = torch.zeros(n * self.P, dtype=torch.float32, device="cuda") if scores else None
        _check(lib().am_plan_infer(ctypes.c_void_p(feat.data_ptr()), n, ctypes.c_void_p(self.w.data_ptr()), self.P,
                                   ctypes.c_void_p(self.mean.data_ptr()), ctypes.c_void_p(self.std.data_ptr()),
                                   ctypes.c_float(self.min_margin), ctypes.c_void_p(fb.data_ptr()),
                                   ctypes.c_void_p(plan.data_ptr()),
                                   ctypes.c_void_p(sc.data_ptr()) if sc is not None else None, _stream()),
               "am_plan_infer")
        torch.cuda.synchronize()
        return plan.cpu().numpy(), (sc.view(n, self.P).cpu().numpy() if sc is not None else None)


def mlp_reference(z, X):
    """numpy forward pass of the same MLP (checks the HIP inference kernel)."""
    w = z["weights"]
    P = int(z["nplans"])
    o = 0
    def take(k):
        nonlocal o
        r = w[o:o + k]
        o += k
        return r
    W1 = take(NHID * NFEAT).reshape(NHID, NFEAT); b1 = take(NHID)
    W2 = take(NHID * NHID).reshape(NHID, NHID); b2 = take(NHID)
    W3 = take(P * NHID).reshape(P, NHID); b3 = take(P)
    x = (X - z["mean"]) / z["std"]
    h = np.maximum(x @ W1.T + b1, 0)
    h = np.maximum(h @ W2.T + b2, 0)
    return h @ W3.T + b3


# ----------------------------------------------------------------- CLI

def _collect(args):
    import amethyst_amd as am
    torch = am._torch()
    X, T, specs = [], [], []
    t0 = time.time()
    for k in range(args.n):
        seed = args.seed + k
        spec = column_spec(seed, nmax=args.nmax)
        a, b, pred = make_inputs(seed, spec)
        f = features([(a, b, pred)]).cpu().numpy()[0]
        T.append(time_plans(a, b, pred, f, reps=args.reps))
        X.append(f)
        specs.append(spec)
        del a, b
        if k % 100 == 0:
            print(f"collect {k}/{args.n} {time.time() - t0:.0f}s", flush=True)
            torch.cuda.empty_cache()
    np.savez(args.out, X=np.array(X, np.float32), T=np.array(T), specs=np.array(specs, dtype=np.int64))
    print("wrote", args.out)


def choose_with_margin(scores, margin):
    """Host restatement of k_plan_mlp's rule: best plan when it beats the
    runner-up by at least `margin`, else the analytic plan."""
    s = np.sort(scores, axis=1)
    arg = scores.argmin(axis=1)
    return np.where(s[:, 1] - s[:, 0] >= margin, arg, ANALYTIC)


def regret(T, choice):
    return float(T[np.arange(len(T)), choice].sum() / T.min(axis=1).sum() - 1)


MARGINS = (0.0, 0.005, 0.01, 0.02, 0.05, 0.1)


def _train(args):
    X, T = [], []
    for path in args.data.split(","):
        z = np.load(path)
        X.append(z["X"])
        T.append(z["T"])
    X, T = np.concatenate(X), np.concatenate(T)
    # hold out 20 % (every 5th column) to choose the confidence margin
    val = np.arange(len(T)) % 5 == 0
    w = train(X[~val], T[~val], epochs=args.epochs)
    sv = mlp_reference(w, X[val])
    res = {m: regret(T[val], choose_with_margin(sv, m)) for m in MARGINS}
    margin = min(res, key=res.get)
    print("validation regret by margin: " + ", ".join(f"{m}: {100 * r:.2f}%" for m, r in res.items()) +
          f"; analytic {100 * regret(T[val], np.full(val.sum(), ANALYTIC)):.2f}%")
    # final fit on all columns, with the chosen margin stored beside the weights
    w = train(X, T, epochs=args.epochs)
    w["margin"] = np.float32(margin)
    np.savez(args.out, **w)
    tr = regret(T, choose_with_margin(mlp_reference(w, X), margin))
    print(f"trained on {len(T)} columns; margin {margin}; training-set regret {100 * tr:.2f}%")


def _evaluate(args):
    import amethyst_amd as am
    torch = am._torch()
    pl = Planner(args.weights, min_margin=args.margin)
    zw = np.load(args.weights)
    rows = []
    t0 = time.time()
    for k in range(args.n):
        seed = args.seed + k
        spec = column_spec(seed, nmax=args.nmax)
        a, b, pred = make_inputs(seed, spec)
        f = features([(a, b, pred)])
        choice, sc = pl.choose(f, scores=True)
        fr = f.cpu().numpy()
        ref = mlp_reference(zw, fr)
        times = time_plans(a, b, pred, fr[0], reps=args.reps)
        rows.append({"spec": list(map(int, spec)), "times": times.tolist(), "choice": int(choice[0]),
                     "mlp_match": bool(np.allclose(sc[0], ref[0], rtol=1e-3, atol=1e-3)),
                     "feat": fr[0].tolist(), "scores": sc[0].tolist()})
        del a, b
        if k % 200 == 0:
            print(f"evaluate {k}/{args.n} {time.time() - t0:.0f}s", flush=True)
            torch.cuda.empty_cache()
    T = np.array([r["times"] for r in rows])
    ch = np.array([r["choice"] for r in rows])
    best = T.min(axis=1)
    S = np.array([r["scores"] for r in rows])
    np.savez(os.path.splitext(args.out)[0] + ".npz", X=np.array([r["feat"] for r in rows], np.float32), T=T,
             S=S, choice=ch, specs=np.array([r["spec"] for r in rows], np.int64))
    res = {
        "columns": len(rows),
        "margin": pl.min_margin,
        "regret_by_margin": {str(m): regret(T, choose_with_margin(S, m)) for m in MARGINS},
        "regret_best_single_plan": float(min(T[:, p].sum() for p in range(NPLANS)) / best.sum() - 1),
        "regret_learned": float(T[np.arange(len(T)), ch].sum() / best.sum() - 1),
        "regret_analytic": float(T[:, ANALYTIC].sum() / best.sum() - 1),
        "mean_col_regret_learned": float(np.mean(T[np.arange(len(T)), ch] / best - 1)),
        "fraction_best": float(np.mean(T[np.arange(len(T)), ch] <= best * 1.0001)),
        "fallback_rate": float(np.mean(ch == ANALYTIC)),
        "hip_mlp_matches_numpy": bool(all(r["mlp_match"] for r in rows)),
        "plans": [f"{'slab' if f == SLAB else 'bitstream'}/{wr}/J{v}" for f, wr, v in PLANS],
        "best_plan_histogram": np.bincount(T.argmin(axis=1), minlength=NPLANS).tolist(),
        "chosen_plan_histogram": np.bincount(ch, minlength=NPLANS).tolist(),
    }
    json.dump(res, open(args.out, "w"), indent=1)
    print(json.dumps(res))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--n", type=int, default=3000)
    c.add_argument("--seed", type=int, default=1000)
    c.add_argument("--nmax", type=int, default=1 << 26)
    c.add_argument("--reps", type=int, default=3)
    c.add_argument("--out", default="plan_train.npz")
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, help="one or more .npz files, comma-separated")
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--out", default=WEIGHTS)
    e = sub.add_parser("evaluate")
    e.add_argument("--n", type=int, default=10000)
    e.add_argument("--seed", type=int, default=900000)
    e.add_argument("--nmax", type=int, default=1 << 26)
    e.add_argument("--reps", type=int, default=2)
    e.add_argument("--margin", type=float, default=None)
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--out", default="plan_eval.json")
    args = ap.parse_args(argv)
    {"collect": _collect, "train": _train, "evaluate": _evaluate}[args.cmd](args)


if __name__ == "__main__":
    main(sys.argv[1:])
