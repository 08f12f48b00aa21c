"""Trains the learned scheduler on the GPU and writes its weights.

    python lambdavm-rust_amd/train_scheduler.py [--n 12 --m 12] [--train-seeds 64]
        [--also radix_lsd,batch,stress,stress_18,bitonic,church,sum,radix] [--also-seeds 8]
        [--eval-runs 12] [--out lambdavm-rust_amd/sched_weights.json]

Workload: church_loop(n, m) of programs.py (an outer loop of 2^n iterations,
each applying the Church numeral 2^m to a u24 increment), evaluated at seeds
0..train_seeds-1 once under each of the 4 actions held fixed for every turn;
`--also` adds other workloads the same way at seeds 0..also_seeds-1
(radix_lsd: the LSD radix sort of 2^18 keys; batch: the 256-net batch of
programs.batch(0); stress, stress_18: the reference's stress program with 2^16
loops of 256 and 2^18 loops of 1024; bitonic, radix, sum, church: the earlier
sort, sum and Church-fan programs), so that one network serves them all.  Every recorded turn of a
workload is a training input, and its targets are the whole-evaluation
speedups of the 4 actions on that workload (log2 of median time under
action 0 over median time under the action).  The MLP is fitted with the
library's HIP kernels; the margin by which it must beat action 0 is picked on
training seeds (the best geometric-mean speedup among margins that keep every
workload within 2% of plain in the median and within 5% at every seed); then
plain and learned schedules are timed on alternating
runs at held-out seeds (which perturb placement and turn lengths in both
modes) on every workload and on programs the network never saw, and both
must give the same normal form and ITRS.  The weights file holds the fp16
parameters (as integers) and the report.
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
    ap.add_argument("--epochs", type=int, default=800)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--eval-runs", type=int, default=12)
    ap.add_argument("--also", default="radix_lsd,batch,stress,stress_18,bitonic,church,sum,radix",
                    help="more workloads to train on: radix_lsd,batch,stress,stress_18,bitonic,radix,sum,church")
    ap.add_argument("--also-seeds", type=int, default=8)
    ap.add_argument("--unseen", default="church_loop_11,stress_14",
                    help="programs evaluated but never trained on: church_loop_11 (church_loop(11, 12)), "
                         "stress_14 (stress with 2^14 loops of 256)")
    ap.add_argument("--margins", type=lambda v: [int(x) for x in v.split(",")], default=[0, 2, 4, 8, 12],
                    help="candidate margins (hundredths of log2 speedup) for the plain fallback")
    ap.add_argument("--margin-seeds", type=int, default=8, help="training seeds each margin is timed on")
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

    def program(name):
        """(label, code, expected result or None, run, context config)"""
        if name == "batch":
            code, nets = P.batch(0, 256)
            fids = P.def_fids(code)
            entries = [fids[f"net{i}"] for i in range(len(nets))]
            est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
            return "batch(0, 256 nets)", code, None, batch_run(entries, est), {"node_cap": 1 << 29, "vars_cap": 1 << 29}
        gens = {"bitonic": ("bitonic(16, 0)", lambda: P.bitonic(16, 0)),
                "radix": ("radix(18, 0)", lambda: P.radix(18, 0)),
                "radix_lsd": ("radix_lsd(18, 0)", lambda: P.radix_lsd(18, 0)),
                "sum": ("sum_tree(20)", lambda: P.sum_tree(20)),
                "church": ("church(12, 12)", lambda: P.church(12, 12)),
                "stress": ("stress(16, 256)", lambda: P.stress(16, 256)),
                "stress_14": ("stress(14, 256)", lambda: P.stress(14, 256)),
                "stress_18": ("stress(18, 1024)", lambda: P.stress(18, 1024)),
                "church_loop_11": ("church_loop(11, 12)", lambda: P.church_loop(11, 12))}
        label, gen = gens[name]
        code, expect = gen()
        # the sorts print trees; the others a number
        return label, code, None if name in ("bitonic", "radix") else str(expect), single, {}

    church = P.church_loop(args.n, args.m)
    progs = {f"church_loop(n={args.n}, m={args.m})": (church[0], str(church[1]), args.train_seeds, single, {})}
    for name in [x for x in args.also.split(",") if x]:
        label, code, expect, run, cfg = program(name)
        progs[label] = (code, expect, args.also_seeds, run, cfg)
    unseen = {}
    for name in [x for x in args.unseen.split(",") if x]:
        label, code, expect, run, cfg = program(name)
        unseen[label] = (code, expect, 0, run, cfg)
    books = {label: H.compile_book(v[0]) for label, v in list(progs.items()) + list(unseen.items())}

    t0 = time.time()
    rows, fixed_ms = [], {}
    for label, (_, _, nseeds, run, cfg) in progs.items():
        runs = S.collect_fixed(books[label], range(nseeds), run=lambda c, run=run: run(c)[0], time_limit_s=300, **cfg)
        r, med = S.fixed_targets(runs)
        fixed_ms[label] = med
        print(f"fixed: {label}: median ms per action {[round(m, 2) for m in med]} ({len(r)} rows, {nseeds} seeds)",
              flush=True)
        rows += r
    print(f"fixed: {len(rows)} rows in {time.time() - t0:.1f}s", flush=True)
    p16, p32, loss, off = S.train(rows, epochs=args.epochs, lr=args.lr)
    print(f"trained: loss {loss:.5f} (target offset {off:.4f})", flush=True)

    def ab(label, spec, seeds, margin, check=False):
        """Plain and learned on alternating runs; returns (plain ms list, learned ms list, identical results)."""
        code, expect, _, run, cfg = spec
        plain, learned, keys = [], [], set()
        with H.Context(time_limit_s=300, flags=0 if check else H.NO_READBACK, **cfg) as ctx:
            ctx.load(books[label])
            run(ctx)
            for i, sd in enumerate(seeds):
                for mode in ((H.SCHED_PLAIN, H.SCHED_LEARNED) if i % 2 == 0 else (H.SCHED_LEARNED, H.SCHED_PLAIN)):
                    ctx.set_scheduler(mode, seed=sd, explore_pct=margin if mode == H.SCHED_LEARNED else 0,
                                      params_fp16=p16 if mode == H.SCHED_LEARNED else None)
                    ms, key, result = run(ctx)
                    (plain if mode == H.SCHED_PLAIN else learned).append(ms)
                    if check:
                        keys.add(key)
                        if expect is not None and result != expect:
                            raise SystemExit(f"{label}: wrong result {result!r}, expected {expect}")
        return plain, learned, len(keys) <= 1

    # The margin by which the network must beat action 0 (hundredths of log2
    # speedup) is chosen on training seeds: the best geometric mean of
    # plain/learned median time over the workloads among the margins that
    # keep every workload within 2% of plain in the median and within 5% at
    # every seed (a schedule can go wrong at one placement only: the Church
    # fan's last turns may run a long sequential fold alone); the largest
    # margin if none does.
    margins, worst, seed_worst = {}, {}, {}
    for m in args.margins:
        rs, sw = [], []
        for label, spec in progs.items():
            pt, lt, _ = ab(label, spec, range(args.margin_seeds), m)
            rs.append(statistics.median(pt) / statistics.median(lt))
            sw.append(min(p / q for p, q in zip(pt, lt)))   # pt[i], lt[i]: seed i
        margins[m] = math.exp(statistics.fmean(math.log(r) for r in rs))
        worst[m] = min(rs)
        seed_worst[m] = min(sw)
        print(f"margin {m}: geo-mean speedup on training seeds {margins[m]:.4f} {[round(r, 3) for r in rs]}, "
              f"worst seed {seed_worst[m]:.3f}", flush=True)
    safe = [m for m in margins if worst[m] >= 0.98 and seed_worst[m] >= 0.95]
    margin = max(safe, key=margins.get) if safe else max(margins)

    evals = {}
    for label, spec in list(progs.items()) + list(unseen.items()):
        first = spec[2]  # held out: no training run used these seeds
        seeds = range(first, first + args.eval_runs)
        plain, learned, same = ab(label, spec, seeds, margin, check=True)
        if not same:
            raise SystemExit(f"{label}: plain and learned results differ")
        evals[label] = {"plain_ms_median": statistics.median(plain),
                        "learned_ms_median": statistics.median(learned),
                        "speedup": statistics.median(plain) / statistics.median(learned),
                        "identical_results": same, "trained_on": label in progs,
                        "eval_seeds": [first, first + args.eval_runs - 1],
                        "fixed_action_ms": fixed_ms.get(label)}
        print(label, json.dumps(evals[label]), flush=True)
    first = next(iter(progs))
    report = dict(evals[first], workload=first, train_rows=len(rows), train_seeds=args.train_seeds,
                  loss=loss, eval_runs=args.eval_runs, trained_on=list(progs), per_workload=evals,
                  margin_pct=margin, margin_search={str(k): v for k, v in margins.items()},
                  margin_worst_seed={str(k): v for k, v in seed_worst.items()})
    print(json.dumps({k: v for k, v in report.items() if k != "per_workload"}), flush=True)
    with open(args.out, "w") as f:
        json.dump({"params_fp16": p16, "reward_offset": off, "margin_pct": margin, "report": report}, f)


if __name__ == "__main__":
    main()
