"""Evaluates one synthetic workload once on the GPU and prints its statistics.

    python lambdavm-rust_amd/run_workload.py stress --args 8,4096 [--mode compiled]

Workloads are the generators of programs.py (sum_tree, bitonic, radix,
radix_lsd, church, church_loop, stress, tree_gen) with their parameters in
order; modes are the
evaluators of include/hvm_hip.h (interpreted, compiled CALLs, local nets,
64-bit ports).  Prints one JSON line: ITRS, evaluation time, MIPS, wave-steps,
turns, spills, and the turn kernel's HBM roofline fraction: the algorithmic
bytes of the interactions (counted by the kernel per interaction kind) per
dispatch over the average dispatch time (HIP events on the launch stream),
against 8 TB/s.  With --repeat N the net is evaluated N times and the fastest
evaluation is reported.  With --phases the turn kernel times its step phases
(HVM_HIP_PROFILE, device wall clock) and the line adds the average wave-step's
µs per phase: refill+share, classify, takes+alloc, creation, links+count,
release wait.  Used for profiles of single workloads.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import hvm_hip as H  # noqa: E402
import programs as P  # noqa: E402

HBM_PEAK_GBS = 8000.0
MODES = {"interpreted": 0, "compiled": H.COMPILED_CALLS, "local": H.LOCAL_NET, "wide": H.WIDE}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("sum_tree", "bitonic", "radix", "radix_lsd", "church", "church_loop",
                                         "stress", "tree_gen"))
    ap.add_argument("--args", default="", help="comma-separated generator parameters")
    ap.add_argument("--mode", choices=tuple(MODES), default="interpreted")
    ap.add_argument("--time-limit", type=int, default=60)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--phases", action="store_true", help="time the step phases (HVM_HIP_PROFILE)")
    a = ap.parse_args()
    params = [int(x) for x in a.args.split(",") if x]
    code = getattr(P, a.workload)(*params)[0]
    words = H.compile_book(code)
    flags = H.NO_READBACK | MODES[a.mode] | (H.PROFILE if a.phases else 0)
    with H.Context(flags=flags, time_limit_s=a.time_limit) as ctx:
        ctx.load(words)
        runs = [ctx.eval() for _ in range(max(1, a.repeat))]
    assert all(x.itrs == runs[0].itrs for x in runs), "ITRS changed between evaluations"
    r = min(runs, key=lambda x: x.eval_ms)
    # the 64-bit-port turn loop carries no events: its dispatches are timed by the wall clock
    launch_ms = r.launch_ms if r.launch_ms > 0 else r.eval_ms
    gbs = r.bytes / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    print(json.dumps({"workload": a.workload, "args": params, "mode": a.mode, "itrs": r.itrs,
                      "eval_ms": round(r.eval_ms, 3), "mips": round(r.itrs / r.eval_ms / 1e3, 1),
                      "wave_steps": r.diag["steps"], "turns": r.turns, "spills": r.diag["spills"],
                      "waves": r.diag["waves"], "launches": r.launches,
                      "bytes_per_launch": round(r.bytes / max(1, r.launches)),
                      "avg_launch_ms": round(launch_ms / max(1, r.launches), 4),
                      "hbm_gbs": round(gbs, 2), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                      **({"us_per_wave_step": [round(ph * 1e3 / max(1, r.diag["steps"]), 3)
                                               for ph in r.diag["phase_ms"]]} if a.phases else {})}), flush=True)


if __name__ == "__main__":
    main()
