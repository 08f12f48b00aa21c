"""Trains the learned turn-length scheduler on the GPU and writes its weights.

    python lambdavm-rust_amd/train_scheduler.py [--n 12 --m 12] [--train-seeds 64]
        [--also bitonic,radix,sum,batch] [--also-seeds 16]
        [--eval-runs 16] [--out lambdavm-rust_amd/sched_weights.json]

Workload: church(n, m) of programs.py (2^n iterations of the Church numeral
2^m applied to a u24 increment).  Traces come from the plain heuristic with
seeded random turn lengths mixed in (seeds 0..train_seeds-1); `--also` adds
traces of other workloads (bitonic 2^16, radix 2^18, sum 2^20, the
256-net batch of programs.batch(0); seeds 0..also_seeds-1), so that one
network serves them all.  The MLP is fitted
with the library's HIP kernels; then plain and learned schedules are timed
on alternating runs at held-out seeds (no trace run is reused) on every
workload, and both must give the same normal form and ITRS.  The weights file
holds the fp16 parameters (as integers), the reward offset and the report.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import hvm_hip as H  # noqa: E402
import programs as P  # noqa: E402
import sched as S  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=12)
    ap.add_argument("--m", type=int, default=12)
    ap.add_argument("--train-seeds", type=int, default=64)
    ap.add_argument("--explore-pct", type=int, default=35)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--eval-runs", type=int, default=16)
    ap.add_argument("--also", default="", help="more workloads to train on: bitonic,radix,sum,batch")
    ap.add_argument("--also-seeds", type=int, default=16)
    ap.add_argument("--margins", type=lambda v: [int(x) for x in v.split(",")], default=[0, 5, 10, 20, 40],
                    help="candidate margins (hundredths of log2 rate) for the plain fallback")
    ap.add_argument("--out", default=os.path.join(HERE, "sched_weights.json"))
    args = ap.parse_args()

    # run(ctx) evaluates a workload: (eval ms, (ITRS, normal form(s)), printed result)
    def single(ctx):
        r = ctx.eval()
        return r.eval_ms, (r.itrs, r.canonical), r.result

    def batch_run(entries, est):
        def run(ctx):
            per, tot = ctx.eval_batch(entries, est)
            return tot.eval_ms, (tot.itrs, tuple(p.canonical for p in per)), None
        return run

    church = P.church(args.n, args.m)
    progs = {f"church(n={args.n}, m={args.m})": (church[0], str(church[1]), args.train_seeds, single, {})}
    extra = {"bitonic": ("bitonic(16, 0)", lambda: P.bitonic(16, 0)),
             "radix": ("radix(18, 0)", lambda: P.radix(18, 0)),
             "sum": ("sum_tree(20)", lambda: P.sum_tree(20))}
    for name in [x for x in args.also.split(",") if x]:
        if name == "batch":
            code, nets = P.batch(0, 256)
            fids = P.def_fids(code)
            entries = [fids[f"net{i}"] for i in range(len(nets))]
            est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
            progs["batch(0, 256 nets)"] = (code, None, args.also_seeds, batch_run(entries, est),
                                           {"node_cap": 1 << 29, "vars_cap": 1 << 29})
            continue
        label, gen = extra[name]
        code, expect = gen()
        progs[label] = (code, None, args.also_seeds, single, {})
    books = {label: H.compile_book(v[0]) for label, v in progs.items()}

    t0 = time.time()
    recs = []
    for label, (_, _, nseeds, run, cfg) in progs.items():
        r = S.collect_traces(books[label], range(nseeds), args.explore_pct, run=run, time_limit_s=300, **cfg)
        print(f"traces: {label}: {len(r)} decisions from {nseeds} runs", flush=True)
        recs += r
    print(f"traces: {len(recs)} decisions in {time.time() - t0:.1f}s", flush=True)
    p16, p32, loss, off = S.train(recs, epochs=args.epochs, lr=args.lr)
    print(f"trained: loss {loss:.4f} (reward offset {off:.3f})", flush=True)

    # The margin by which the network must beat the plain heuristic's choice
    # (hundredths of log2 rate) is chosen on training seeds: geometric mean of
    # plain/learned time over the workloads.
    def ratio(label, seeds, margin):
        pt, lt = [], []
        run, cfg = progs[label][3], progs[label][4]
        with H.Context(time_limit_s=300, flags=H.NO_READBACK, **cfg) as ctx:
            ctx.load(books[label])
            run(ctx)
            for sd in seeds:
                ctx.set_scheduler(H.SCHED_PLAIN, seed=sd)
                pt.append(run(ctx)[0])
                ctx.set_scheduler(H.SCHED_LEARNED, seed=sd, explore_pct=margin, params_fp16=p16)
                lt.append(run(ctx)[0])
        return statistics.median(pt) / statistics.median(lt)
    margins = {}
    for m in args.margins:
        rs = [ratio(label, range(4), m) for label in progs]
        margins[m] = math.exp(statistics.fmean(math.log(r) for r in rs))
        print(f"margin {m}: geo-mean speedup on training seeds {margins[m]:.4f} {[round(r, 3) for r in rs]}", flush=True)
    margin = max(margins, key=margins.get)

    evals = {}
    for label, (code, expect, nseeds, run, cfg) in progs.items():
        plain, learned, results = [], [], set()
        with H.Context(time_limit_s=300, **cfg) as ctx:
            ctx.load(books[label])
            for i in range(args.eval_runs):
                seed = nseeds + i  # held out: no trace run used it
                for mode in ((H.SCHED_PLAIN, H.SCHED_LEARNED) if i % 2 == 0 else (H.SCHED_LEARNED, H.SCHED_PLAIN)):
                    ctx.set_scheduler(mode, seed=seed, explore_pct=margin if mode == H.SCHED_LEARNED else 0,
                                      params_fp16=p16 if mode == H.SCHED_LEARNED else None)
                    ms, key, result = run(ctx)
                    (plain if mode == H.SCHED_PLAIN else learned).append(ms)
                    results.add(key)
                    if expect is not None and result != expect:
                        raise SystemExit(f"wrong result {result!r}, expected {expect}")
        evals[label] = {"plain_ms_median": statistics.median(plain),
                        "learned_ms_median": statistics.median(learned),
                        "speedup": statistics.median(plain) / statistics.median(learned),
                        "identical_results": len(results) == 1,
                        "eval_seeds": [nseeds, nseeds + args.eval_runs - 1]}
        print(label, json.dumps(evals[label]), flush=True)
    first = next(iter(progs))
    report = dict(evals[first], workload=first, train_decisions=len(recs), train_seeds=args.train_seeds,
                  loss=loss, eval_runs=args.eval_runs, trained_on=list(progs), per_workload=evals,
                  margin_pct=margin, margin_search={str(k): v for k, v in margins.items()})
    print(json.dumps({k: v for k, v in report.items() if k != "per_workload"}), flush=True)
    with open(args.out, "w") as f:
        json.dump({"params_fp16": p16, "reward_offset": off, "margin_pct": margin, "report": report}, f)


if __name__ == "__main__":
    main()
