"""Benchmark: MIPS of the HIP evaluator on one MI355X.

Workload (one step): full normalisation of a bitonic sort of 2^16 u24 keys
drawn from SplitMix64(seed 0), stored as a perfect depth-16 tree — the program
bitonic(16, 0) of lambdavm-rust_amd/programs.py.  The book, net and bags are
resident in HBM before timing; each step clears the net, boots @main and runs
the evaluator to normal form (with the IO driver's readback expansions).
Steps are timed with the wall clock around K steps; the dominant kernel
(hvm_turn_kernel) is timed with HIP events on its own stream inside the
library.  Parity (ITRS and normal form vs the reference run-c golden record)
is checked on a warmup step outside the timed region.  The timed steps
interpret the book as run-cu does; the compiled-CALL mode (def templates
pre-reduced at load, same ITRS) and the 64-bit-port evaluator are checked
and timed after them and reported in "mode_compare" (`--mode` picks which
of the three the timed steps use).

Prints one JSON line (see the contract in the repository instructions).
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lambdavm-rust_amd"))

METRIC = "MIPS (interactions/sec) learned vs plain sched, 1 GPU, bit-exact NF; %HBM roofline"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(words, itrs_expected):
    """Reference run-c (oracle/_ref) on this host's cores, else the C port."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    ncpu = min(ncpu, 16)  # the GPU box's CPU share
    threads = 1
    while threads * 2 <= ncpu and threads * 2 <= 32:
        threads *= 2
    if oracle.have_reference(threads):
        r = oracle.run_reference(words, threads, timeout=300)
        if r.itrs == itrs_expected and r.time_s:
            return {"value": round(r.itrs / r.time_s / 1e6, 2), "unit": "MIPS", "cores": threads,
                    "kind": "reference",
                    "sample": f"one full run of the bench workload by the reference run-c "
                              f"(src/run.c, -O3 -DIO, TPC={threads}); time window boot→result as run-c prints it"}
    o = oracle.run_oracle(words)
    return {"value": round(o.itrs / o.seconds / 1e6, 2), "unit": "MIPS", "cores": 1, "kind": "port",
            "sample": "one full run of the bench workload by the single-thread C restatement (oracle/)"}


def profile_values(k):
    """key: value lines of the committed counter summary for this workload."""
    path = os.path.join(ROOT, "profiles", f"hvm_turn_kernel.bitonic{k}.md")
    vals = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.strip().strip("`").split(":")
            if len(parts) == 2 and parts[0].strip().isidentifier():
                try:
                    vals[parts[0].strip()] = float(parts[1])
                except ValueError:
                    pass
    return vals, os.path.relpath(path, ROOT)


def atomic_roofline(k, itrs_per_s):
    """The evaluator's exchange rate against the measured device-scope
    exchange rate on random slots (hvm_hip_atomic_rate): a 64 MiB buffer (fits
    the 256 MiB Infinity Cache, as the bench net's live part does) and a
    1 GiB one (HBM)."""
    import hvm_hip
    vals, src = profile_values(k)
    per_itr = vals.get("TCC_ATOMIC_per_interaction")
    peaks = {f"{b}B_{name}": hvm_hip.atomic_rate(slots, b) / 1e9
             for b in (4, 8) for name, slots in (("64MiB", (64 << 20) // b), ("1GiB", (1 << 30) // b))}
    peak = max(peaks["4B_64MiB"], peaks["8B_64MiB"])
    out = {"unit": "Gatomic/s", "peak": round(peak, 2),
           "peaks_measured": {k2: round(v, 2) for k2, v in peaks.items()}}
    if per_itr:
        achieved = itrs_per_s * per_itr / 1e9
        out.update({"achieved": round(achieved, 2), "frac": round(achieved / peak, 4),
                    "atomics_per_interaction": per_itr, "source": src})
    return out


def measured_traffic(k):
    """HBM bytes per turn-kernel dispatch for this workload, from the committed
    rocprofv3 counter summary: FETCH_SIZE and WRITE_SIZE (KiB, summed over the
    profiled dispatches), each scaled by the factor the same counter needs to
    read a plain device copy of known size correctly.  None if absent."""
    path = os.path.join(ROOT, "profiles", f"hvm_turn_kernel.bitonic{k}.md")
    if not os.path.exists(path):
        return None, None
    vals = {}
    for line in open(path):
        parts = line.strip().strip("`").split(":")
        if len(parts) == 2 and parts[0].strip().isidentifier():
            try:
                vals[parts[0].strip()] = float(parts[1])
            except ValueError:
                pass
    need = ("FETCH_SIZE_KiB", "WRITE_SIZE_KiB", "fetch_copy_factor", "write_copy_factor", "turn_dispatches")
    if not all(n in vals for n in need):
        return None, None
    b = (vals["FETCH_SIZE_KiB"] * vals["fetch_copy_factor"] +
         vals["WRITE_SIZE_KiB"] * vals["write_copy_factor"]) * 1024 / vals["turn_dispatches"]
    return round(b, 1), os.path.relpath(path, ROOT)


def radix_line(hvm_hip, P, golden, flags):
    """The radix sort of 2^18 u24 keys (SplitMix64 seed 0, 24 one-bit passes),
    evaluated after the timed region in the bench's mode: the fastest of 3
    evaluations, its MIPS and the turn kernel's HBM roofline fraction
    (algorithmic bytes per dispatch over the average dispatch time), with ITRS
    and normal form checked against the golden record."""
    code, _ = P.radix(18, 0)
    words = hvm_hip.compile_book(code)
    g = next((c for c in golden["workloads"]
              if c["generator"] == "radix" and c["params"] == {"k": 18, "seed": 0}), None)
    with hvm_hip.Context(time_limit_s=600, flags=flags) as c:
        c.load(words)
        rv = c.eval()
    with hvm_hip.Context(time_limit_s=600, flags=hvm_hip.NO_READBACK | flags) as c:
        c.load(words)
        rs = [c.eval() for _ in range(3)]
    assert all(r.itrs == rv.itrs for r in rs), "ITRS changed between evaluations"
    r = min(rs, key=lambda x: x.eval_ms)
    launch_ms = r.launch_ms if r.launch_ms > 0 else r.eval_ms
    gbs = r.bytes / (launch_ms * 1e-3) / 1e9
    return {"workload": "radix sort of 2^18 u24 keys (SplitMix64 seed 0), 24 one-bit passes",
            "mips": round(r.itrs / r.eval_ms / 1e3, 2), "ms": round(r.eval_ms, 3),
            "achieved_gbs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "itrs_match": g is not None and rv.itrs == g["itrs"],
            "nf_match": g is not None and
            hashlib.sha256(rv.canonical.encode()).hexdigest() == g["nf_sha256"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=16, help="log2 of the key count")
    ap.add_argument("--sched", choices=("plain", "learned"), default="plain",
                    help="turn-length scheduler of the timed steps (the other one is timed after them)")
    ap.add_argument("--no-sched-compare", action="store_true",
                    help="do not time the other scheduler (profiling runs: only the bench's own evaluations)")
    ap.add_argument("--mode", choices=("interpreted", "compiled", "wide"), default="interpreted",
                    help="evaluator of the timed steps: the book interpreted as run-cu does (default), "
                         "compiled CALLs (templates pre-reduced at load, fast CALL paths), or the "
                         "64-bit-port evaluator; the other two are checked and timed after them")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (torch), to calibrate "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE against a known byte count")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this evaluator runs a net on a single device)", file=sys.stderr)
        sys.exit(2)

    import hvm_hip
    import programs as P

    code, sorted_keys = P.bitonic(args.k, 0)
    words = hvm_hip.compile_book(code)
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    gcase = next((c for c in golden["workloads"]
                  if c["generator"] == "bitonic" and c["params"] == {"k": args.k, "seed": 0}), None)

    MODES = {"interpreted": 0, "compiled": hvm_hip.COMPILED_CALLS, "wide": hvm_hip.WIDE}
    mode = MODES[args.mode]
    wide = args.mode == "wide"
    ctx = hvm_hip.Context(time_limit_s=600, flags=mode)
    ctx.load(words)
    # parity step (untimed): full readback, compared with the golden record
    r0 = ctx.eval()
    parity = {"itrs_match": gcase is not None and r0.itrs == gcase["itrs"],
              "nf_match": gcase is not None and
              hashlib.sha256(r0.canonical.encode()).hexdigest() == gcase["nf_sha256"],
              "sorted_keys_match": P.flatten_tree(P.parse_tree_text(r0.result)) == sorted_keys}
    ctx.close()

    weights, margin = None, 0
    wpath = os.path.join(ROOT, "lambdavm-rust_amd", "sched_weights.json")
    if os.path.exists(wpath):
        wj = json.load(open(wpath))
        weights, margin = wj["params_fp16"], wj.get("margin_pct", 0)
    if args.sched == "learned" and weights is None:
        sys.exit("bench.py: --sched learned needs lambdavm-rust_amd/sched_weights.json")

    def use(ctx, sched):
        if sched == "learned":
            ctx.set_scheduler(hvm_hip.SCHED_LEARNED, explore_pct=margin, params_fp16=weights)
        else:
            ctx.set_scheduler(hvm_hip.SCHED_PLAIN)

    ctx = hvm_hip.Context(time_limit_s=600, flags=hvm_hip.NO_READBACK | mode)
    ctx.load(words)
    if not wide:
        use(ctx, args.sched)
    for _ in range(max(0, args.warmup - 1)):
        ctx.eval()
    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runs.append(ctx.eval())
    t1 = time.perf_counter()
    # the other scheduler on the same workload, after the timed region
    other = "plain" if args.sched == "learned" else "learned"
    compare = {args.sched + "_mips": None, other + "_mips": None}
    if weights is not None and not args.no_sched_compare and not wide:
        use(ctx, other)
        ctx.eval()
        t2 = time.perf_counter()
        o_runs = [ctx.eval() for _ in range(args.steps)]
        t3 = time.perf_counter()
        assert all(r.itrs == runs[0].itrs for r in o_runs), "ITRS changed with the scheduler"
        compare[other + "_mips"] = round(runs[0].itrs * args.steps / (t3 - t2) / 1e6, 2)
    ctx.close()

    # the other evaluation modes (interpreted / compiled CALLs, and the
    # 64-bit-port evaluator), after the timed region: parity against the
    # golden record, then the same step count
    variant = []
    if not args.no_sched_compare:
        for name, vflag in MODES.items():
            if name == args.mode:
                continue
            with hvm_hip.Context(time_limit_s=600, flags=vflag) as vctx:
                vctx.load(words)
                rv = vctx.eval()
            with hvm_hip.Context(time_limit_s=600, flags=hvm_hip.NO_READBACK | vflag) as vctx:
                vctx.load(words)
                if not vflag & hvm_hip.WIDE:
                    use(vctx, args.sched)
                vctx.eval()
                t4 = time.perf_counter()
                v_runs = [vctx.eval() for _ in range(args.steps)]
                t5 = time.perf_counter()
            variant.append({"mode": name,
                            "value": round(v_runs[0].itrs * args.steps / (t5 - t4) / 1e6, 2),
                            "ms_per_step": round((t5 - t4) / args.steps * 1e3, 3),
                            "itrs_match": gcase is not None and rv.itrs == gcase["itrs"],
                            "nf_match": gcase is not None and
                            hashlib.sha256(rv.canonical.encode()).hexdigest() == gcase["nf_sha256"]})

    itrs = runs[0].itrs
    assert all(r.itrs == itrs for r in runs), "ITRS changed between steps"
    total_s = t1 - t0
    ms_per_step = total_s / args.steps * 1e3
    value = itrs * args.steps / total_s / 1e6
    kernel_ms = sum(r.kernel_ms for r in runs)
    # every dispatch of the turn kernel, including those that found the net
    # already normal, as rocprofv3 lists them
    launches = sum(r.launches for r in runs)
    launch_ms = sum(r.launch_ms for r in runs)
    if wide:
        # the wide turn loop is timed by the wall clock (its kernels carry no events)
        kernel_ms = launch_ms = sum(r.eval_ms for r in runs)
    bytes_total = sum(r.bytes for r in runs)
    avg_launch_ms = launch_ms / launches
    bytes_per_launch = bytes_total / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    # the counter profile is of the interpreting kernel
    traffic, traffic_src = measured_traffic(args.k) if args.mode == "interpreted" else (None, None)
    kernel_name = {"interpreted": "hvm_turn_kernel", "compiled": "hvm_turn_fast_kernel",
                   "wide": "hvm_wide_turn_kernel"}[args.mode]
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MIPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"bitonic sort of 2^{args.k} u24 keys (SplitMix64 seed 0), depth-{args.k} key tree",
                   "keys": 1 << args.k, "itrs_per_step": itrs, "scheduler": args.sched,
                   "mode": args.mode},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": kernel_name, "launches_per_step": launches / args.steps,
                     "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
                     "avg_launch_ms": round(avg_launch_ms, 4)},
        "kernel_mips": round(itrs * args.steps / (kernel_ms * 1e-3) / 1e6, 2),
        "eval_ms_per_step": round(sum(r.eval_ms for r in runs) / args.steps, 3),
        "parity": parity,
        "schedulers": compare,
        "mode_compare": variant,
    }
    compare[args.sched + "_mips"] = out["value"]
    if args.mode == "interpreted":
        out["atomic_roofline"] = atomic_roofline(args.k, itrs * args.steps / (kernel_ms * 1e-3))
    if not args.no_sched_compare:
        out["radix18"] = radix_line(hvm_hip, P, golden, mode)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(words, itrs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
