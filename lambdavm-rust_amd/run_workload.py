"""Evaluates one synthetic workload once on the GPU and prints its statistics.

    python lambdavm-rust_amd/run_workload.py stress --args 8,4096 [--mode compiled]

Workloads are the generators of programs.py (sum_tree, bitonic, radix,
church, stress, tree_gen) with their parameters in order; modes are the
evaluators of include/hvm_hip.h (interpreted, compiled CALLs, local nets,
64-bit ports).  Prints one JSON line: ITRS, evaluation time, MIPS, wave-steps,
turns, spills.  Used for profiles of single workloads.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import hvm_hip as H  # noqa: E402
import programs as P  # noqa: E402

MODES = {"interpreted": 0, "compiled": H.COMPILED_CALLS, "local": H.LOCAL_NET, "wide": H.WIDE}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("sum_tree", "bitonic", "radix", "church", "stress", "tree_gen"))
    ap.add_argument("--args", default="", help="comma-separated generator parameters")
    ap.add_argument("--mode", choices=tuple(MODES), default="interpreted")
    ap.add_argument("--time-limit", type=int, default=60)
    a = ap.parse_args()
    params = [int(x) for x in a.args.split(",") if x]
    code = getattr(P, a.workload)(*params)[0]
    words = H.compile_book(code)
    with H.Context(flags=H.NO_READBACK | MODES[a.mode], time_limit_s=a.time_limit) as ctx:
        ctx.load(words)
        r = ctx.eval()
    print(json.dumps({"workload": a.workload, "args": params, "mode": a.mode, "itrs": r.itrs,
                      "eval_ms": round(r.eval_ms, 3), "mips": round(r.itrs / r.eval_ms / 1e3, 1),
                      "wave_steps": r.diag["steps"], "turns": r.turns, "spills": r.diag["spills"],
                      "waves": r.diag["waves"]}), flush=True)


if __name__ == "__main__":
    main()
