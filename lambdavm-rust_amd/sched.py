"""Learned scheduler: trace collection, training and the A/B test.

After every turn the evaluator's scheduler chooses how the next turn shares
work, one of 4 (wave share, idle-quit share, steady bound) actions: (all,
the configured 25%, 1), (all, 50%, 1), (half, 25%, 1), (all, 25%, 16).  A turn
runs on that share of the waves and ends once its idle-quit share of them has
run dry; then the HBM bag redistributes the redexes.  After a turn that did
not grow the redexes and left at most (steady bound)/4 per wave, the next
turn runs without idle quit (steady chains) until a wave holds surplus work.
The plain heuristic always takes action 0.  The learned
scheduler is a 32->64->64->4 MLP (fp16 weights, fp32 accumulation) that reads
the finished turn's features — pending redexes, lane occupancy, per-kind
interaction mix, spill/defer/probe rates, how the turn ended (waves run dry,
steady, growth), per-wave occupancy at its end (waves that left 0, 1-64,
65-128, >128 redexes), free-slot density, rate — and predicts, for each
action, log2 of the speedup of the whole evaluation when that action is
taken; it takes the argmax, and keeps action 0 unless the prediction beats
it by a margin.

* `collect_fixed` evaluates a program once per (seed, action) with the action
  fixed for every turn (fixed mode), recording each turn's features and the
  evaluation's time;
* `fixed_targets` turns those runs into training rows: every recorded turn
  of a program gets, for each action, log2(median time under action 0 /
  median time under that action) — the reward is the whole evaluation's
  time, not the next turn's;
* `train` fits the MLP on those rows with the library's HIP training kernels
  (full-batch Adam);
* `ab_test` times plain vs learned and checks that both give the same normal
  form and interaction count.

The scheduler never changes an interaction, so results are bit-identical
whichever policy runs; only time changes.
"""
from __future__ import annotations

import math
import random
import statistics
from typing import Dict, List, Sequence, Tuple

import hvm_hip as H


def collect_fixed(words, seeds: Sequence[int], run=None, **cfg) -> Dict[int, List[Tuple[float, list]]]:
    """Evaluates the program once per seed and action, the action fixed for
    every turn (`run(ctx)` evaluates it and returns its time in ms; default
    ctx.eval()).  Actions alternate within a seed, so that slow drifts of the
    device hit all of them alike.  Returns {action: [(ms, [features per
    turn]), ...]}."""
    run = run or (lambda c: c.eval().eval_ms)
    out: Dict[int, List[Tuple[float, list]]] = {a: [] for a in range(H.MLP_ACT)}
    with H.Context(flags=H.NO_READBACK, **cfg) as ctx:
        ctx.load(words)
        run(ctx)
        for s in seeds:
            for a in (range(H.MLP_ACT) if s % 2 == 0 else reversed(range(H.MLP_ACT))):
                ctx.set_scheduler(H.SCHED_FIXED, seed=s, explore_pct=a, trace_cap=1 << 16)
                ms = run(ctx)
                out[a].append((ms, [feats for feats, action, reward, itrs, us in ctx.read_trace()]))
    return out


def fixed_targets(runs: Dict[int, List[Tuple[float, list]]],
                  per_run: int = 64) -> Tuple[List[Tuple[list, int, float]], List[float]]:
    """Training rows of one program: every turn recorded under any action
    (at most `per_run` turns of a run, evenly spaced) gets one row per action
    a, with target log2(T0 / Ta), where Ta is the median evaluation time with
    a fixed.  Returns (rows, [Ta])."""
    med = [statistics.median(ms for ms, _ in runs[a]) for a in range(len(runs))]
    gain = [math.log2(med[0] / m) for m in med]
    rows = []
    for a in range(len(runs)):
        for _, turns in runs[a]:
            step = max(1.0, len(turns) / per_run)
            for f in (turns[int(i * step)] for i in range(min(per_run, len(turns)))):
                rows += [(f, k, gain[k]) for k in range(len(runs))]
    return rows, med


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
