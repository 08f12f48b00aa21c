"""Benchmark: MIPS of the HIP evaluator on one MI355X.

Workload (one step): full normalisation of an LSD radix sort of 2^18 u24 keys
drawn uniform (duplicates kept) from SplitMix64(seed 0), one bit per pass
over 24 passes, each pass a tree map plus a stable merge, then the checksum
Σ sorted[i]·i mod 2^24 — the program radix_lsd(18, 0) of
lambdavm-rust_amd/programs.py.  The book, net and bags are resident in HBM
before timing; each step clears the net, boots @main, runs the evaluator to
normal form and reads the result back (the window run-c reports).  Steps are
timed with the wall clock around K steps after W warmup steps; the dominant
kernel (hvm_turn_kernel) is timed with HIP events on its own stream inside
the library.  Parity (ITRS and normal form vs the reference run-c golden
record) is checked on the first warmup step, and ITRS on every timed step.

Both schedulers are timed with the same protocol (W warmup + K steps each):
the plain heuristic and the learned MLP; `value` is the scheduler chosen by
--sched (learned by default).  The 256-net batched workload is timed the same
way for both schedulers.  After the timed regions: the bitonic sort of 2^16
keys (with its own roofline) and the trie radix sort of 2^18 distinct keys.

Prints one JSON line (see the contract in the repository instructions).
"""
import argparse
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lambdavm-rust_amd"))

METRIC = "MIPS (interactions/sec) learned vs plain sched, 1 GPU, bit-exact NF; %HBM roofline"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, params, description, profile summary)
    "radix_lsd18": ("radix_lsd", {"k": 18, "seed": 0},
                    "LSD radix sort of 2^18 uniform u24 keys (SplitMix64 seed 0, duplicates kept), "
                    "24 one-bit passes of tree map + stable merge, checksum sum(sorted[i]*i) mod 2^24",
                    "hvm_turn_kernel.radix_lsd18.md"),
    "bitonic16": ("bitonic", {"k": 16, "seed": 0},
                  "bitonic sort of 2^16 u24 keys (SplitMix64 seed 0), depth-16 key tree",
                  "hvm_turn_kernel.bitonic16.md"),
    "radix18": ("radix", {"k": 18, "seed": 0},
                "trie radix sort of 2^18 distinct u24 keys (SplitMix64 seed 0)",
                "hvm_turn_kernel.radix18.md"),
}


def cpu_baseline(P, golden):
    """The reference run-c (oracle/_ref, src/run.c compiled in place) on this
    host's cores: median of 5 runs of a bounded sample of the bench workload,
    the LSD radix sort of 2^16 keys (same program, a quarter of the keys);
    else the single-thread C port."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hvm_hip
    import oracle
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = min(nproc, 16)  # the GPU box's CPU share
    threads = 1
    while threads * 2 <= share and threads * 2 <= 32:
        threads *= 2
    code, expect = P.radix_lsd(16, 0)
    words = hvm_hip.compile_book(code)
    sample = "radix_lsd(16, 0): LSD radix sort of 2^16 uniform u24 keys, 24 passes (a bounded sample of the bench workload)"
    if oracle.have_reference(threads):
        mips, itrs = [], None
        for _ in range(5):
            r = oracle.run_reference(words, threads, timeout=300)
            if r.result != str(expect) or not r.time_s:
                break
            itrs = r.itrs
            mips.append(r.itrs / r.time_s / 1e6)
        if len(mips) == 5:
            return {"value": round(statistics.median(mips), 2), "unit": "MIPS", "cores": threads, "kind": "reference",
                    "nproc": nproc, "tpc": threads, "runs_mips": [round(m, 2) for m in mips], "itrs": itrs,
                    "sample": sample + f"; median of 5 runs of the reference run-c (src/run.c -O3 -DIO, TPC={threads}, "
                                       f"nproc={nproc}), time window boot->result as run-c prints it"}
    o = oracle.run_oracle(words)
    return {"value": round(o.itrs / o.seconds / 1e6, 2), "unit": "MIPS", "cores": 1, "kind": "port",
            "nproc": nproc, "tpc": 1, "sample": sample + "; one run of the single-thread C restatement (oracle/)"}


def profile_values(name):
    """key: value lines of the committed rocprofv3 summary for a workload."""
    path = os.path.join(ROOT, "profiles", name)
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


def measured_traffic(name):
    """HBM bytes per turn-kernel dispatch, from the committed counter summary:
    FETCH_SIZE and WRITE_SIZE (KiB, summed over the profiled dispatches), each
    scaled by the factor the same counter needs to read a plain device copy
    of known size correctly.  None if absent."""
    vals, src = profile_values(name)
    need = ("FETCH_SIZE_KiB", "WRITE_SIZE_KiB", "fetch_copy_factor", "write_copy_factor", "turn_dispatches")
    if not all(n in vals for n in need):
        return None, None
    b = (vals["FETCH_SIZE_KiB"] * vals["fetch_copy_factor"] +
         vals["WRITE_SIZE_KiB"] * vals["write_copy_factor"]) * 1024 / vals["turn_dispatches"]
    return round(b, 1), src


def atomic_roofline(name, itrs_per_s):
    """The evaluator's memory-side atomic rate against the measured
    device-scope exchange rate on random slots (hvm_hip_atomic_rate): a 64 MiB
    buffer (fits the 256 MiB Infinity Cache, as the live part of the net does)
    and a 1 GiB one."""
    import hvm_hip
    vals, src = profile_values(name)
    per_itr = vals.get("TCC_ATOMIC_per_interaction")
    peaks = {f"{b}B_{n}": hvm_hip.atomic_rate(slots, b) / 1e9
             for b in (4, 8) for n, slots in (("64MiB", (64 << 20) // b), ("1GiB", (1 << 30) // b))}
    peak = max(peaks["4B_64MiB"], peaks["8B_64MiB"])
    out = {"unit": "Gatomic/s", "peak": round(peak, 2), "peaks_measured": {k: round(v, 2) for k, v in peaks.items()}}
    if per_itr:
        achieved = itrs_per_s * per_itr / 1e9
        out.update({"achieved": round(achieved, 2), "frac": round(achieved / peak, 4),
                    "atomics_per_interaction": per_itr, "source": src})
    return out


def request_roofline(name, itrs_per_s):
    """The evaluator against its memory-side request ceiling.  Per
    interaction it issues E exchanges (TCC_EA0_ATOMIC), S other writes
    (TCC_EA0_WRREQ, which counts the atomics too, minus E) and R reads
    (TCC_EA0_RDREQ), from the committed counter summary; the device sustains
    rE exchanges, rS device-scope stores and rR device-scope loads per second
    on random slots of a 64 MiB buffer (hvm_hip_request_rate, 8-byte slots).
    With those requests as the only cost an interaction takes E/rE + S/rS +
    R/rR seconds of the memory side: the peak interaction rate."""
    import hvm_hip
    vals, src = profile_values(name)
    e, w, r = (vals.get(k) for k in ("TCC_ATOMIC_per_interaction", "TCC_EA0_WRREQ_per_interaction",
                                     "TCC_EA0_RDREQ_per_interaction"))
    if not (e and w and r):
        return None
    slots = (64 << 20) // 8
    rates = {k: hvm_hip.request_rate(slots, 8, op) for k, op in (("exchange", 0), ("store", 1), ("load", 2))}
    mix = {"exchange": e, "store": w - e, "load": r}
    peak = 1.0 / sum(mix[k] / rates[k] for k in mix)
    return {"unit": "G interactions/s", "peak": round(peak / 1e9, 3), "achieved": round(itrs_per_s / 1e9, 3),
            "frac": round(itrs_per_s / peak, 4),
            "requests_per_interaction": {k: round(v, 3) for k, v in mix.items()},
            "rates_measured_G_per_s": {k: round(v / 1e9, 2) for k, v in rates.items()}, "source": src}


def stream_copy_gbs(mib=1024):
    """Measured device copy bandwidth (bytes read plus written per second) of
    a `mib` MiB buffer, best of 5 device-to-device copies
    (hvm_hip_copy_rate): the attainable HBM peak beside the 8 TB/s
    specification."""
    import hvm_hip
    return hvm_hip.copy_rate(mib << 20) / 1e9


def golden_case(golden, gen, params):
    return next((c for c in golden["workloads"] if c["generator"] == gen and c["params"] == params), None)


def tuned_run(name, golden):
    """One autotuned configuration (lambdavm-rust_amd/tuned_<name>.json),
    evaluated by `autotune.py --child` in a child process that loads the
    configuration's build: MIPS and ms per evaluation, and whether ITRS and
    normal form(s) equal the run-c records; None without the file."""
    import subprocess
    path = os.path.join(ROOT, "lambdavm-rust_amd", f"tuned_{name}.json")
    if not os.path.exists(path):
        return None
    cfg = json.load(open(path))["config"]
    env = dict(os.environ)
    if cfg.get("variant"):
        env["HVM_HIP_LIB"] = os.path.join(ROOT, "lambdavm-rust_amd", "lib", "variants",
                                          f"libhvm_hip_{cfg['variant']}.so")
    p = subprocess.run([sys.executable, os.path.join(ROOT, "lambdavm-rust_amd", "autotune.py"), "--child",
                        "--workload", name, "--cfg", json.dumps(cfg), "--repeats", "3", "--readback",
                        "--time-limit", "300"], capture_output=True, text=True, env=env, timeout=900)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    if "error" in res:
        raise RuntimeError(f"tuned configuration of {name} failed: {res['error']}")
    if name == "batch":
        gb = json.load(open(os.path.join(ROOT, "tests", "golden", "batch.json")))
        itrs = gb["total_itrs"]
        ok = res["ident"][0] == itrs and res["nets_nf_sha256"] == [row["nf_sha256"] for row in gb["nets"]]
    else:
        case = next(c for c in golden["workloads"] if c["generator"] == name and c.get("full_size"))
        itrs = case["itrs"]
        ok = res["ident"] == [case["itrs"], case["nf_sha256"]]
    return {"config": cfg, "mips": round(itrs / (res["ms"] * 1e-3) / 1e6, 2), "ms_per_eval": round(res["ms"], 3),
            "parity": bool(ok), "protocol": "median of 3 evaluations after a warm-up (single nets: readback "
                                            "included; the batch: timed without it, normal forms read back "
                                            "from one more evaluation)"}


def timed(ctx, steps, warmup, run):
    """W warmup evaluations, then exactly K timed ones (device synchronised
    on both sides: every evaluation ends with the result read back)."""
    for _ in range(warmup):
        run(ctx)
    t0 = time.perf_counter()
    rs = [run(ctx) for _ in range(steps)]
    t1 = time.perf_counter()
    return rs, t1 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="radix_lsd18")
    ap.add_argument("--sched", choices=("plain", "learned"), default="learned",
                    help="scheduler whose timed steps are the headline value (both are timed)")
    ap.add_argument("--mode", choices=("interpreted", "compiled"), default="interpreted",
                    help="the book interpreted as run-cu does (default) or with compiled CALLs")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the headline scheduler only (profiling runs: the bench's own evaluations)")
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

    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    gen, params, desc, prof = WORKLOADS[args.workload]
    code, expect = getattr(P, gen)(**params)
    words = hvm_hip.compile_book(code)
    gcase = golden_case(golden, gen, params)
    mode = hvm_hip.COMPILED_CALLS if args.mode == "compiled" else 0

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

    def check(r):
        if gen == "bitonic":
            return P.flatten_tree(P.parse_tree_text(r.result)) == expect
        return r.result == str(expect)

    # ---- headline: the chosen scheduler, then the other, same protocol ----
    ctx = hvm_hip.Context(time_limit_s=600, flags=mode)
    ctx.load(words)
    r0 = ctx.eval()
    parity = {"itrs_match": gcase is not None and r0.itrs == gcase["itrs"],
              "nf_match": gcase is not None and
              hashlib.sha256(r0.canonical.encode()).hexdigest() == gcase["nf_sha256"],
              "closed_form_match": check(r0)}
    scheds = [args.sched] + ([] if args.only_headline or weights is None else
                             ["learned" if args.sched == "plain" else "plain"])
    per_sched = {}
    runs = None
    for sname in scheds:
        use(ctx, sname)
        rs, secs = timed(ctx, args.steps, args.warmup, lambda c: c.eval())
        assert all(r.itrs == r0.itrs for r in rs), "ITRS changed between steps"
        assert all(check(r) for r in rs), "result changed between steps"
        per_sched[sname] = {"mips": round(r0.itrs * args.steps / secs / 1e6, 2),
                            "ms_per_step": round(secs / args.steps * 1e3, 3)}
        if sname == args.sched:
            runs, total_s = rs, secs
    ctx.close()
    if len(per_sched) == 2:
        per_sched["learned_over_plain"] = round(per_sched["learned"]["mips"] / per_sched["plain"]["mips"], 4)

    itrs = runs[0].itrs
    value = itrs * args.steps / total_s / 1e6
    kernel_ms = sum(r.kernel_ms for r in runs)
    # every dispatch of the turn kernel, including those that found the net
    # already normal, as rocprofv3 lists them
    launches = sum(r.launches for r in runs)
    launch_ms = sum(r.launch_ms for r in runs)
    bytes_total = sum(r.bytes for r in runs)
    avg_launch_ms = launch_ms / launches
    bytes_per_launch = bytes_total / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(prof) if args.mode == "interpreted" else (None, None)
    kernel_name = "hvm_turn_fast_kernel" if args.mode == "compiled" else "hvm_turn_kernel"
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MIPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": desc, "program": f"{gen}({', '.join(str(v) for v in params.values())})",
                   "itrs_per_step": itrs, "scheduler": args.sched, "mode": args.mode},
        "roofline": {"bound": "hbm", "limiter": "memory-side requests and their round trips (see request_roofline)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src, "kernel": kernel_name,
                     "launches_per_step": launches / args.steps,
                     "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
                     "avg_launch_ms": round(avg_launch_ms, 4)},
        "kernel_mips": round(itrs * args.steps / (kernel_ms * 1e-3) / 1e6, 2),
        "eval_ms_per_step": round(sum(r.eval_ms for r in runs) / args.steps, 3),
        "parity": parity,
        "schedulers": {args.workload: per_sched},
    }
    copy_gbs = stream_copy_gbs()
    out["roofline"]["stream_copy_gbs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_stream_copy"] = round(achieved / copy_gbs, 4)
    if args.mode == "interpreted":
        out["atomic_roofline"] = atomic_roofline(prof, itrs * args.steps / (kernel_ms * 1e-3))
        out["request_roofline"] = request_roofline(prof, itrs * args.steps / (kernel_ms * 1e-3))

    if not args.only_headline:
        # ---- the 256-net batch, both schedulers, same protocol ----
        gb = json.load(open(os.path.join(ROOT, "tests", "golden", "batch.json")))
        bcode, nets = P.batch(gb["seed"], gb["n"])
        fids = P.def_fids(bcode)
        entries = [fids[f"net{i}"] for i in range(len(nets))]
        est = [P.batch_estimate(g, p) for g, p, _, _ in nets]
        bsched = {}
        with hvm_hip.Context(time_limit_s=600, node_cap=1 << 29, vars_cap=1 << 29,
                             flags=mode | hvm_hip.NO_READBACK) as bctx:
            bctx.load(hvm_hip.compile_book(bcode))
            for sname in scheds:
                use(bctx, sname)
                rs, secs = timed(bctx, args.steps, max(1, args.warmup // 2), lambda c: c.eval_batch(entries, est)[1])
                assert all(r.itrs == gb["total_itrs"] for r in rs), "batch ITRS differs from the run-c record"
                bsched[sname] = {"mips": round(gb["total_itrs"] * args.steps / secs / 1e6, 2),
                                 "ms_per_step": round(secs / args.steps * 1e3, 3)}
        if len(bsched) == 2:
            bsched["learned_over_plain"] = round(bsched["learned"]["mips"] / bsched["plain"]["mips"], 4)
        out["schedulers"]["batch256"] = bsched

        # ---- the autotuner's configurations (lambdavm-rust_amd/tuned_<name>.json:
        # compiled CALLs, idle-quit share, turn length, graphs, compile-time
        # variant), each measured by the autotuner's own protocol (median of 3
        # evaluations after one warm-up) in a child process that loads that
        # build, ITRS and NF checked against the run-c records ----
        tuned = {}
        for key, tname in (("radix_lsd18", "radix_lsd"), ("batch256", "batch")):
            if key != "batch256" and key != args.workload:
                continue
            t = tuned_run(tname, golden)
            if t is not None:
                tuned[key] = t
        out["tuned"] = tuned

        # ---- the other sort workloads, after the timed regions ----
        extra = {}
        for wname, (g2, p2, d2, prof2) in WORKLOADS.items():
            if wname == args.workload:
                continue
            c2, e2 = getattr(P, g2)(**p2)
            w2 = hvm_hip.compile_book(c2)
            gc2 = golden_case(golden, g2, p2)
            # the normal form is read back from one untimed evaluation; the
            # timed ones skip the readback (a sort prints 2^16-2^18 keys)
            with hvm_hip.Context(time_limit_s=600, flags=mode) as c:
                c.load(w2)
                rv = c.eval()
            with hvm_hip.Context(time_limit_s=600, flags=mode | hvm_hip.NO_READBACK) as c:
                c.load(w2)
                c.eval()
                rs, secs = timed(c, 3, 0, lambda cc: cc.eval())
            assert all(r.itrs == rv.itrs for r in rs)
            la = sum(r.launches for r in rs)
            lms = sum(r.launch_ms for r in rs)
            gbs = sum(r.bytes for r in rs) / la / (lms / la * 1e-3) / 1e9
            extra[wname] = {"workload": d2, "mips": round(rv.itrs * 3 / secs / 1e6, 2),
                            "ms_per_step": round(secs / 3 * 1e3, 3), "roofline_gbs": round(gbs, 2),
                            "frac": round(gbs / HBM_PEAK_GBS, 4),
                            "itrs_match": gc2 is not None and rv.itrs == gc2["itrs"],
                            "nf_match": gc2 is not None and
                            hashlib.sha256(rv.canonical.encode()).hexdigest() == gc2["nf_sha256"]}
        out["other_workloads"] = extra
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, golden)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
