"""Trains the learned turn-length scheduler on the GPU and writes its weights.

    python lambdavm-rust_amd/train_scheduler.py [--n 12 --m 12] [--train-seeds 64]
        [--eval-runs 16] [--out lambdavm-rust_amd/sched_weights.json]

Workload: church(n, m) of programs.py (2^n iterations of the Church numeral
2^m applied to a u24 increment).  Traces come from the plain heuristic with
seeded random turn lengths mixed in (seeds 0..train_seeds-1); the MLP is fitted
with the library's HIP kernels; then plain and learned schedules are timed
on alternating runs (held-out: no trace run is reused) and both must give the
same normal form and ITRS.  The weights file holds the fp16 parameters (as
integers), the reward offset and the A/B report.
"""
import argparse
import json
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
    ap.add_argument("--out", default=os.path.join(HERE, "sched_weights.json"))
    args = ap.parse_args()

    code, expect = P.church(args.n, args.m)
    words = H.compile_book(code)
    t0 = time.time()
    recs = S.collect_traces(words, range(args.train_seeds), args.explore_pct, time_limit_s=300)
    t1 = time.time()
    print(f"traces: {len(recs)} decisions from {args.train_seeds} runs in {t1 - t0:.1f}s", flush=True)
    p16, p32, loss, off = S.train(recs, epochs=args.epochs, lr=args.lr)
    print(f"trained: loss {loss:.4f} (reward offset {off:.3f})", flush=True)

    plain, learned, results = [], [], set()
    with H.Context(time_limit_s=300) as ctx:
        ctx.load(words)
        for i in range(args.eval_runs):
            seed = args.train_seeds + i  # held out: no trace run used it
            for mode in ((H.SCHED_PLAIN, H.SCHED_LEARNED) if i % 2 == 0 else (H.SCHED_LEARNED, H.SCHED_PLAIN)):
                ctx.set_scheduler(mode, seed=seed, params_fp16=p16 if mode == H.SCHED_LEARNED else None)
                r = ctx.eval()
                (plain if mode == H.SCHED_PLAIN else learned).append(r.eval_ms)
                results.add((r.itrs, r.canonical))
                if r.result != str(expect):
                    raise SystemExit(f"wrong result {r.result!r}, expected {expect}")
    report = {
        "workload": f"church(n={args.n}, m={args.m})",
        "train_decisions": len(recs), "train_seeds": args.train_seeds, "loss": loss,
        "plain_ms_median": statistics.median(plain), "learned_ms_median": statistics.median(learned),
        "speedup": statistics.median(plain) / statistics.median(learned),
        "identical_results": len(results) == 1,
        "eval_runs": args.eval_runs,
        "eval_seeds": [args.train_seeds, args.train_seeds + args.eval_runs - 1],
    }
    print(json.dumps(report), flush=True)
    with open(args.out, "w") as f:
        json.dump({"params_fp16": p16, "reward_offset": off, "report": report}, f)


if __name__ == "__main__":
    main()
