"""Autotuner for the evaluator's launch and scheduling knobs.

    python lambdavm-rust_amd/autotune.py [--workload batch|bitonic|radix|sum|church]
        [--repeats 2] [--out lambdavm-rust_amd/tuned_<workload>.json]

The knobs are the ones a caller can set per context (include/hvm_hip.h,
hvm_hip_cfg):
  * waves        — evaluator workgroups (one 64-lane wave each) per turn, as
                   a fraction of the default, which fills every CU as far as
                   registers and LDS allow;
  * idle_pct     — a turn ends once this % of the waves ran out of work, so
                   that the output bag redistributes the busy waves' redexes;
  * turn_steps   — a fixed share interval (steps per turn) instead of the
                   scheduler's choice;
  * flags        — HVM_HIP_LOCAL_NET (per-wave LDS nets) and/or
                   HVM_HIP_COMPILED_CALLS (def templates pre-reduced at load);
  * graphs       — HVM_HIP_GRAPHS (turn chunks replayed from HIP graphs) or
                   not, or-ed into the flags.
The per-lane redex deque (HVM_DEQ) and the LDS net sizes are compile-time
constants of the kernels and are not searched here.

Search: coordinate descent from the defaults, one knob at a time, keeping a
value only when its median time beats the best so far.  Every candidate must
reproduce the workload's interaction count (ITRS) and, for single nets, the
canonical normal form; a candidate that does not is an error, not a slow
point, and so is an evaluator error flag (it is raised, not skipped); only
candidates that the library rejects at context creation or book load, or
that run out of the time limit, are recorded as not runnable.  The result is
written as JSON: the chosen configuration, the default
and tuned times, and every measurement.
"""
from __future__ import annotations

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


class Workload:
    """A book plus how to evaluate it once on a context."""

    def __init__(self, name, words, runner, itrs=None):
        self.name, self.words, self.runner, self.itrs = name, words, runner, itrs


def single(name: str, code: str) -> Workload:
    words = H.compile_book(code)
    return Workload(name, words, lambda ctx: ctx.eval())


def batched(seed: int = 0, n: int = 256) -> Workload:
    code, nets = P.batch(seed, n)
    fids = P.def_fids(code)
    entries = [fids[f"net{i}"] for i in range(len(nets))]
    est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
    words = H.compile_book(code)
    return Workload(f"batch(seed={seed}, n={n})", words, lambda ctx: ctx.eval_batch(entries, est)[1])


def measure(w: Workload, cfg: dict, repeats: int, time_limit_s: int = 60):
    """Median eval time (ms) of `repeats` runs after one warm-up; also the
    identity (ITRS, canonical NF) of the result."""
    flags = cfg.get("flags", 0) | cfg.get("graphs", 0) | (H.NO_READBACK if w.name.startswith("batch") else 0)
    ctx_cfg = {k: v for k, v in cfg.items() if k != "graphs"}
    with H.Context(time_limit_s=time_limit_s, **dict(ctx_cfg, flags=flags)) as ctx:
        ctx.load(w.words)
        first = w.runner(ctx)
        ms = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            r = w.runner(ctx)
            ms.append((time.perf_counter() - t0) * 1e3)
            if r.itrs != first.itrs:
                raise RuntimeError(f"ITRS changed between runs under {cfg}: {first.itrs} vs {r.itrs}")
    ident = (first.itrs, None if w.name.startswith("batch") else first.canonical)
    return statistics.median(ms), ident, first


def tune(w: Workload, repeats: int = 2, log=print, space=None, base_cfg=None):
    """Coordinate descent over the knobs; returns the report dict."""
    base = dict(base_cfg or {})
    t_def, ident, r0 = measure(w, base, repeats)
    waves_def = r0.diag.get("waves", 0)
    log(f"[{w.name}] default: {t_def:.2f} ms, waves {waves_def}, itrs {ident[0]}")
    space = space or [
        # fractions of the default (all resident waves)
        ("waves", [waves_def * k // 8 for k in (2, 4, 6, 7)] + [0]),
        ("idle_pct", [10, 25, 75, 101, 0]),
        ("turn_steps", [96, 512, 4096, 0]),
        ("flags", [H.COMPILED_CALLS, H.LOCAL_NET, H.LOCAL_NET | H.COMPILED_CALLS, 0]),
        ("graphs", [H.GRAPHS, 0]),
    ]
    best, t_best = dict(base), t_def
    rows = [{"cfg": dict(base), "ms": round(t_def, 3)}]
    for knob, values in space:
        for v in values:
            cand = dict(best)
            cand[knob] = v
            if knob == "flags" and v & H.LOCAL_NET:
                cand["waves"] = 0   # its own resident-wave count
            if cand == best:
                continue
            try:
                t, ident2, _ = measure(w, cand, repeats)
            except H.HvmEvalError as e:
                # an evaluator error flag is a failure of the evaluator, never
                # a slow point; only running out of the time limit is skipped
                if e.err or "time limit" not in str(e):
                    raise
                log(f"  {knob}={v}: over the time limit ({e})")
                rows.append({"cfg": dict(cand), "error": str(e)})
                continue
            except H.HvmError as e:
                # rejected at context creation or book load: e.g. a net too
                # large for the local-net evaluator's 2^28 cap
                log(f"  {knob}={v}: not runnable ({e})")
                rows.append({"cfg": dict(cand), "error": str(e)})
                continue
            if ident2 != ident:
                raise RuntimeError(f"result changed under {cand}: {ident2[0]} vs {ident[0]}")
            rows.append({"cfg": dict(cand), "ms": round(t, 3)})
            log(f"  {knob}={v}: {t:.2f} ms")
            if t < t_best:
                best, t_best = cand, t
    log(f"[{w.name}] tuned: {best} {t_best:.2f} ms ({t_def / t_best:.3f}x)")
    return {"workload": w.name, "itrs": ident[0], "default_ms": round(t_def, 3),
            "tuned_ms": round(t_best, 3), "speedup": round(t_def / t_best, 4),
            "config": best, "measurements": rows, "repeats": repeats}


def workload(name: str):
    """The workload and the configuration the search starts from."""
    if name == "batch":
        # 256 nets of up to 2^22 leaves need more than the default 2^28 slots
        return batched(), {"node_cap": 1 << 29, "vars_cap": 1 << 29}
    if name == "bitonic":
        return single("bitonic(16, 0)", P.bitonic(16, 0)[0]), {}
    if name == "radix":
        return single("radix(18, 0)", P.radix(18, 0)[0]), {}
    if name == "sum":
        return single("sum_tree(20)", P.sum_tree(20)[0]), {}
    if name == "church":
        return single("church(12, 12)", P.church(12, 12)[0]), {}
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batch")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    w, base = workload(args.workload)
    rep = tune(w, repeats=args.repeats, log=lambda s: print(s, flush=True), base_cfg=base)
    out = args.out or os.path.join(HERE, f"tuned_{args.workload}.json")
    with open(out, "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps({k: rep[k] for k in ("workload", "default_ms", "tuned_ms", "speedup", "config")}), flush=True)


if __name__ == "__main__":
    main()
