"""Autotuner for the evaluator's launch and scheduling knobs.

    python lambdavm-rust_amd/autotune.py [--workload batch|radix_lsd|church_loop|bitonic|radix|sum|church]
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
                   not, or-ed into the flags;
and the turn kernel's compile-time knobs, through variant builds of the
library (lib/variants/libhvm_hip_<name>.so, `make -C lambdavm-rust_amd
variants`), each measured in a child process that loads it (HVM_HIP_LIB):
  * variant      — deq16 / deq64: per-lane redex deque of 16 / 64 entries
                   instead of 32 (LDS per wave, hence waves per CU);
                   scan4 / scan32: slice windows an allocation scans before
                   the shared pool (4 / 32 instead of 8); lnet512 / lnet2048:
                   local-net sizes of the local-net evaluator (1024 default),
                   tried only when that evaluator is selected; pool: the
                   mid-turn redex pool (HVM_MID_POOL).
The kernel runs one wave per workgroup; waves per workgroup is not a knob.

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
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import hvm_hip as H  # noqa: E402
import programs as P  # noqa: E402


class Workload:
    """A book plus how to evaluate it once on a context; `key` names it for
    workload() (child processes rebuild it from the key)."""

    def __init__(self, name, words, runner, itrs=None, key=None):
        self.name, self.words, self.runner, self.itrs, self.key = name, words, runner, itrs, key


def single(name: str, code: str) -> Workload:
    words = H.compile_book(code)
    return Workload(name, words, lambda ctx: ctx.eval())


def batched(seed: int = 0, n: int = 256) -> Workload:
    code, nets = P.batch(seed, n)
    fids = P.def_fids(code)
    entries = [fids[f"net{i}"] for i in range(len(nets))]
    est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
    words = H.compile_book(code)
    w = Workload(f"batch(seed={seed}, n={n})", words, None)

    def run(ctx):
        per, tot = ctx.eval_batch(entries, est)
        w.last_per = per
        return tot
    w.runner = run
    return w


VARIANTS = ("deq16", "deq64", "scan4", "scan32", "lnet512", "lnet2048", "pool")


def variant_lib(name: str) -> str:
    return os.path.join(HERE, "lib", "variants", f"libhvm_hip_{name}.so")


def measure(w: Workload, cfg: dict, repeats: int, time_limit_s: int = 60):
    """Median eval time (ms) of `repeats` runs after one warm-up; also the
    identity (ITRS, canonical NF) of the result.  A variant build is measured
    in a child process that loads that library."""
    if cfg.get("variant"):
        return measure_child(w, cfg, repeats, time_limit_s)
    return measure_here(w, cfg, repeats, time_limit_s)


def measure_child(w: Workload, cfg: dict, repeats: int, time_limit_s: int = 60):
    lib = variant_lib(cfg["variant"])
    if not os.path.exists(lib):
        raise H.HvmError(f"{lib} is missing: build it with `make -C lambdavm-rust_amd variants`")
    env = dict(os.environ, HVM_HIP_LIB=lib)
    p = subprocess.run([sys.executable, os.path.abspath(__file__), "--child", "--workload", w.key,
                        "--cfg", json.dumps(cfg), "--repeats", str(repeats), "--time-limit", str(time_limit_s)],
                       capture_output=True, text=True, env=env, timeout=time_limit_s * (repeats + 2) + 120)
    line = (p.stdout.strip().splitlines() or [""])[-1]
    try:
        res = json.loads(line)
    except ValueError:
        raise RuntimeError(f"variant child failed under {cfg}: {p.stderr[-800:]}")
    if "error" in res:
        raise (H.HvmEvalError(res["error"], res.get("err", 0)) if res.get("eval") else H.HvmError(res["error"]))
    return res["ms"], tuple(res["ident"]), None


def measure_here(w: Workload, cfg: dict, repeats: int, time_limit_s: int = 60, readback: bool = False):
    flags = cfg.get("flags", 0) | cfg.get("graphs", 0) | \
        (H.NO_READBACK if w.name.startswith("batch") and not readback else 0)
    ctx_cfg = {k: v for k, v in cfg.items() if k not in ("graphs", "variant")}
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
    ident = (first.itrs, None if w.name.startswith("batch") else
             hashlib.sha256(first.canonical.encode()).hexdigest())
    return (statistics.median(ms) if ms else 0.0), ident, first


def tune(w: Workload, repeats: int = 2, log=print, space=None, base_cfg=None):
    """Coordinate descent over the knobs; returns the report dict."""
    base = dict(base_cfg or {})
    t_def, ident, r0 = measure(w, base, repeats)
    waves_def = r0.diag.get("waves", 0) if r0 is not None else 0
    log(f"[{w.name}] default: {t_def:.2f} ms, waves {waves_def}, itrs {ident[0]}")
    space = space or [
        # fractions of the default (all resident waves)
        ("waves", [waves_def * k // 8 for k in (2, 4, 6, 7)] + [0]),
        ("idle_pct", [10, 25, 75, 101, 0]),
        ("turn_steps", [96, 512, 4096, 0]),
        ("flags", [H.COMPILED_CALLS, H.LOCAL_NET, H.LOCAL_NET | H.COMPILED_CALLS, 0]),
        ("graphs", [H.GRAPHS, 0]),
        ("variant", ["deq16", "deq64", "scan4", "scan32", "lnet512", "lnet2048", "pool", ""]),
    ]
    best, t_best = dict(base), t_def
    rows = [{"cfg": dict(base), "ms": round(t_def, 3)}]
    for knob, values in space:
        for v in values:
            cand = dict(best)
            cand[knob] = v
            if knob == "flags" and v & H.LOCAL_NET:
                cand["waves"] = 0   # its own resident-wave count
            if knob == "variant" and v.startswith("lnet") and not cand.get("flags", 0) & H.LOCAL_NET:
                continue            # local-net sizes matter to the local-net evaluator only
            if not cand.get("variant"):
                cand.pop("variant", None)
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
    w, base = _workload(name)
    w.key = name
    return w, base


def _workload(name: str):
    if name == "batch":
        # 256 nets of up to 2^22 leaves need more than the default 2^28 slots
        return batched(), {"node_cap": 1 << 29, "vars_cap": 1 << 29}
    if name == "radix_lsd":
        return single("radix_lsd(18, 0)", P.radix_lsd(18, 0)[0]), {}
    if name == "church_loop":
        return single("church_loop(12, 12)", P.church_loop(12, 12)[0]), {}
    if name == "bitonic":
        return single("bitonic(16, 0)", P.bitonic(16, 0)[0]), {}
    if name == "radix":
        return single("radix(18, 0)", P.radix(18, 0)[0]), {}
    if name == "sum":
        return single("sum_tree(20)", P.sum_tree(20)[0]), {}
    if name == "church":
        return single("church(12, 12)", P.church(12, 12)[0]), {}
    raise SystemExit(f"unknown workload {name}")


def child(args):
    """One measurement in this process (the library named by HVM_HIP_LIB);
    prints one JSON line."""
    w, _ = workload(args.workload)
    cfg = json.loads(args.cfg)
    try:
        ms, ident, _ = measure_here(w, cfg, args.repeats, args.time_limit)
        out = {"ms": ms, "ident": list(ident), "lib": H.LIB_PATH}
        if args.readback and w.name.startswith("batch"):
            # timed without the readback of 256 results; one more evaluation
            # reads every net's normal form back
            _, ident2, _ = measure_here(w, cfg, 0, args.time_limit, readback=True)
            if ident2 != ident:
                raise RuntimeError(f"ITRS changed with readback: {ident2[0]} vs {ident[0]}")
            out["nets_nf_sha256"] = [hashlib.sha256((p.canonical or "").encode()).hexdigest() for p in w.last_per]
        print(json.dumps(out), flush=True)
    except H.HvmEvalError as e:
        print(json.dumps({"error": str(e), "err": e.err, "eval": True}), flush=True)
    except H.HvmError as e:
        print(json.dumps({"error": str(e)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batch")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--out", default=None)
    ap.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cfg", default="{}", help=argparse.SUPPRESS)
    ap.add_argument("--time-limit", type=int, default=60, help=argparse.SUPPRESS)
    ap.add_argument("--readback", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.child:
        return child(args)
    w, base = workload(args.workload)
    rep = tune(w, repeats=args.repeats, log=lambda s: print(s, flush=True), base_cfg=base)
    out = args.out or os.path.join(HERE, f"tuned_{args.workload}.json")
    with open(out, "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps({k: rep[k] for k in ("workload", "default_ms", "tuned_ms", "speedup", "config")}), flush=True)


if __name__ == "__main__":
    main()
