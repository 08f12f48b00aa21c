"""Learned turn-length scheduler: trace collection, training and the A/B test.

The evaluator's scheduler chooses, after every turn, how many steps the next
turn runs before the HBM bag redistributes redexes (one of 24/96/512/4096).
The plain heuristic looks only at the number of pending redexes.  The learned
scheduler is a 32->64->64->4 MLP (fp16 weights, fp32 accumulation) that reads
the finished turn's features — pending redexes, lane occupancy, per-kind
interaction mix, spill/defer/probe rates, the previous choice — and predicts,
for each turn length, the log2 interactions-per-µs of the turn it would
govern; it takes the argmax.

* `collect_traces` runs the plain heuristic with seeded random choices mixed
  in (explore mode) and records (features, choice, reward) per decision;
* `train` fits the MLP on those records with the library's HIP training
  kernels (full-batch Adam);
* `ab_test` times plain vs learned on other seeds and checks that both give
  the same normal form and interaction count.

The scheduler never changes an interaction, so results are bit-identical
whichever policy runs; only time changes.
"""
from __future__ import annotations

import math
import random
import statistics
from typing import List, Sequence, Tuple

import hvm_hip as H


def collect_traces(words, seeds: Sequence[int], explore_pct: int = 35, run=None, **cfg):
    """Runs the program once per seed in explore mode (`run(ctx)` evaluates
    it; default ctx.eval()); returns the decision records that have a
    measured reward."""
    recs = []
    with H.Context(flags=H.NO_READBACK, **cfg) as ctx:
        ctx.load(words)
        for s in seeds:
            ctx.set_scheduler(H.SCHED_EXPLORE, seed=s, explore_pct=explore_pct, trace_cap=1 << 16)
            (run or (lambda c: c.eval()))(ctx)
            for feats, action, reward, itrs, us in ctx.read_trace():
                if us > 0:
                    recs.append((feats, action, reward))
    return recs


def init_params(seed: int = 0) -> List[float]:
    """He-style initialisation, packed W1 b1 W2 b2 W3 b3."""
    rng = random.Random(seed)
    F, Hd, K = H.MLP_FEAT, H.MLP_HID, H.MLP_ACT
    p: List[float] = []
    p += [rng.gauss(0, math.sqrt(2.0 / F)) for _ in range(F * Hd)]
    p += [0.0] * Hd
    p += [rng.gauss(0, math.sqrt(2.0 / Hd)) for _ in range(Hd * Hd)]
    p += [0.0] * Hd
    p += [rng.gauss(0, math.sqrt(1.0 / Hd)) for _ in range(Hd * K)]
    p += [0.0] * K
    assert len(p) == H.MLP_PARAMS
    return p


def train(recs, epochs: int = 400, lr: float = 3e-3, seed: int = 0):
    """Fits the MLP to the rewards of the recorded choices.  Returns
    (fp16 params, fp32 params, final loss, reward offset)."""
    x = [r[0] for r in recs]
    a = [r[1] for r in recs]
    rew = [r[2] for r in recs]
    off = statistics.fmean(rew)
    y = [v - off for v in rew]
    p16, p32, loss = H.mlp_train(init_params(seed), x, a, y, epochs=epochs, lr=lr)
    return p16, p32, loss, off


def run_policy(words, mode: int, params16=None, repeats: int = 3, **cfg) -> Tuple[float, H.RunResult]:
    """Median wall time (ms) of full evaluations under one policy."""
    times = []
    last = None
    with H.Context(**cfg) as ctx:
        ctx.load(words)
        ctx.set_scheduler(mode, params_fp16=params16)
        for _ in range(repeats):
            r = ctx.eval()
            times.append(r.eval_ms)
            last = r
    return statistics.median(times), last


def ab_test(words, params16, repeats: int = 5, **cfg):
    """Plain vs learned on the same program: times and result equality."""
    t_plain, r_plain = run_policy(words, H.SCHED_PLAIN, None, repeats, **cfg)
    t_learn, r_learn = run_policy(words, H.SCHED_LEARNED, params16, repeats, **cfg)
    return {
        "plain_ms": t_plain, "learned_ms": t_learn,
        "speedup": t_plain / t_learn if t_learn > 0 else None,
        "itrs_equal": r_plain.itrs == r_learn.itrs,
        "nf_equal": r_plain.canonical == r_learn.canonical,
        "itrs": r_plain.itrs,
    }
