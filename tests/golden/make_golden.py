"""Writes tests/golden/golden.json: expected normal forms and interaction counts.

Run in the build container (needs /root/reference and oracle/_ref built):
    python tests/golden/make_golden.py                  # everything
    python tests/golden/make_golden.py --only radix_lsd,church_loop
                                  # re-record the workloads of these generators only

Each case records the program text (the reference's own test programs and
examples are its test data; the synthetic workloads come from
lambdavm-rust_amd/programs.py), the canonical normal form and ITRS printed by
the compiled reference C runtime (`run-c`, oracle/_ref/hvm_ref_t8), and, for
the reference programs, the normal form held in the reference's insta
snapshot (tests/snapshots/*.snap).  Large normal forms are stored as a SHA-256
of their canonical text.
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "lambdavm-rust_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hvm_hip  # noqa: E402
import oracle  # noqa: E402
import programs as P  # noqa: E402

REF = "/root/reference"
MAX_INLINE = 4096


def snapshot_nf(rel: str):
    """Result line of the reference's snapshot for a program, if any."""
    key = rel.replace("tests/programs/", "").replace("examples/", "").replace("/", "__")
    cands = glob.glob(os.path.join(REF, "tests", "snapshots", f"run__file@{key}.snap"))
    if not cands:
        return None
    text = open(cands[0]).read().split("---", 2)[-1].strip()
    for line in text.splitlines():
        if line.startswith("Result: "):
            return line[len("Result: "):]
    return text


def record(name, code, extra=None, threads=8, timeout=3000):
    words = hvm_hip.compile_book(code)
    r = oracle.run_reference(words, threads, timeout=timeout)
    if r.itrs is None:
        raise RuntimeError(f"{name}: reference produced no ITRS:\n{r.stdout[-500:]}\n{r.stderr[-500:]}")
    canon = hvm_hip.canonical(r.result)
    case = {"name": name, "itrs": r.itrs, "ref_time_s": r.time_s,
            "nf_sha256": hashlib.sha256(canon.encode()).hexdigest(), "nf_len": len(canon)}
    if len(canon) <= MAX_INLINE:
        case["nf"] = canon
    if extra:
        case.update(extra)
    print(f"{name}: itrs={r.itrs} t={r.time_s}s nf_len={len(canon)}", flush=True)
    return case


SMALL = ([("sum_tree", {"depth": d}) for d in (1, 4, 10, 16)]
         + [("bitonic", {"k": k, "seed": s}) for k, s in ((1, 0), (3, 1), (6, 2), (8, 3), (10, 4), (12, 5))]
         + [("radix", {"k": k, "seed": s}) for k, s in ((1, 0), (4, 1), (8, 2), (10, 3), (12, 4))]
         + [("church", {"n": n, "m": m}) for n, m in ((1, 2), (3, 4), (6, 8), (8, 12))]
         + [("radix_lsd", {"k": k, "seed": s}) for k, s in ((1, 0), (4, 1), (8, 2), (10, 3), (12, 4))]
         + [("radix_lsd", {"k": 6, "seed": 5, "bits": 3})]
         + [("church_loop", {"n": n, "m": m}) for n, m in ((1, 2), (3, 4), (6, 8), (8, 12))])
FULL = [("sum_tree", {"depth": 20}), ("bitonic", {"k": 16, "seed": 0}),
        ("radix", {"k": 18, "seed": 0}), ("church", {"n": 12, "m": 12}),
        ("radix_lsd", {"k": 18, "seed": 0}), ("church_loop", {"n": 12, "m": 12}),
        ("stress", {"depth": 18, "length": 1024})]


def workload_case(gen, params):
    code, expect = getattr(P, gen)(**params)
    name = gen + "_" + "_".join(f"{k}{v}" for k, v in params.items())
    case = record(name, code, {"generator": gen, "params": params, "full_size": (gen, params) in FULL})
    case["closed_form"] = "sorted keys" if gen == "bitonic" else str(expect)
    return case


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--only":
        gens = set(sys.argv[2].split(","))
        path = os.path.join(ROOT, "tests", "golden", "golden.json")
        out = json.load(open(path))
        out["workloads"] = [c for c in out["workloads"] if c["generator"] not in gens]
        for gen, params in SMALL + FULL:
            if gen in gens:
                out["workloads"].append(workload_case(gen, params))
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
    out = {"reference": "compiled reference run-c (src/run.c + src/hvm.c, -O3 -DIO -DTPC_L2=3)",
           "programs": [], "workloads": []}
    files = sorted(glob.glob(os.path.join(REF, "examples", "*", "*.hvm")) +
                   glob.glob(os.path.join(REF, "tests", "programs", "**", "*.hvm"), recursive=True))
    for path in files:
        rel = os.path.relpath(path, REF)
        code = open(path).read()
        if "@test-io = 1" in code or "@test-skip = 1" in code or "/io/" in rel or "demo_io" in rel:
            continue
        try:
            hvm_hip.compile_book(code)
        except hvm_hip.HvmError as e:
            out["programs"].append({"name": rel, "code": code, "compile_error": str(e),
                                    "snapshot": snapshot_nf(rel)})
            print(f"{rel}: compile error {e}")
            continue
        snap = snapshot_nf(rel)
        case = record(rel, code, {"code": code, "snapshot": snap})
        out["programs"].append(case)

    # Edge cases: results that are an eraser, a number, a REF (expanded by the
    # IO driver's readback), free variables, the safe-REF copy rule, erasing a
    # large structure, and a long list.
    edges = {
        "edge_era": "@main = *",
        "edge_num": "@main = 42",
        "edge_ref_result": "@main = @f\n@f = (a a)",
        "edge_ref_number": "@main = @g\n@g = r & 1 ~ $([+] $(2 r))",
        "edge_free_vars": "@main = (a (b (b a)))",
        "edge_safe_ref_copy": "@main = (a b) & @id ~ {a b}\n@id = (x x)",
        "edge_unsafe_ref_expand": "@main = (a b) & @d ~ {a b}\n@d = ({x y} (x y))",
        "edge_erase_tree": "@main = r & @mk ~ (12 *) & 7 ~ r\n"
                           "@mk = (?((0 @mk_s) r) r)\n@mk_s = ({a b} (l u)) & @mk ~ (a l) & @mk ~ (b u)",
        "edge_void_pairs": "@main = r & * ~ 1 & @f ~ * & 3 ~ @f & r ~ 5\n@f = *",
        "edge_list_3000": "@main = r & @ls ~ (3000 r)\n@ls = (?((* @ls_s) r) r)\n"
                          "@ls_s = ({a b} (a t)) & @ls ~ (b t)",
    }
    for name, code in edges.items():
        out["programs"].append(record(name, code, {"code": code, "snapshot": None}))

    for gen, params in SMALL + FULL:
        out["workloads"].append(workload_case(gen, params))
    with open(os.path.join(ROOT, "tests", "golden", "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
