"""Writes tests/golden/batch.json: per-net parity targets of the batched workload.

Run in the build container (needs oracle/_ref built):
    python tests/golden/make_golden_batch.py

For every net of programs.batch(seed=0, n=256) the same net, as a program of
its own, is run by the compiled reference run-c (oracle/_ref/hvm_ref_t8); its
canonical normal form (SHA-256) and ITRS are recorded.  The batched run must
reproduce every per-net normal form, and its total ITRS must equal the sum.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "lambdavm-rust_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hvm_hip  # noqa: E402
import oracle  # noqa: E402
import programs as P  # noqa: E402


def one(args):
    i, (gen, params, expect, code) = args
    w = hvm_hip.compile_book(code)
    r = oracle.run_reference(w, 2, timeout=1200)
    canon = hvm_hip.canonical(r.result)
    if gen == "bitonic":
        assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
    else:
        assert r.result == str(expect)
    return {"net": i, "generator": gen, "params": params, "itrs": r.itrs,
            "nf_sha256": hashlib.sha256(canon.encode()).hexdigest()}


def main():
    code, nets = P.batch(0, 256)
    with ThreadPoolExecutor(4) as ex:
        rows = list(ex.map(one, enumerate(nets)))
    out = {"seed": 0, "n": 256, "total_itrs": sum(r["itrs"] for r in rows), "nets": rows}
    with open(os.path.join(ROOT, "tests", "golden", "batch.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("total itrs", out["total_itrs"])


if __name__ == "__main__":
    main()
