"""Parity of the local-net evaluator (HVM_HIP_LOCAL_NET) with the reference.

Same bar as test_gpu_parity.py: canonical normal form and ITRS bit-identical
to run-c (tests/golden/golden.json).  The local-net evaluator allocates in
each wave's LDS net and copies nodes to HBM only through LEAK interactions
(src/hvm.cu:1129-1187) and at the end of every turn (src/hvm.cu:1426-1482);
neither is an interaction of the program, so ITRS must not move.  Short
turns force an evacuation of every local net after each step.
"""
import hashlib

import pytest

import programs as P

pytestmark = pytest.mark.gpu


def sha(s):
    return hashlib.sha256(s.encode()).hexdigest()


@pytest.fixture(scope="module")
def lctx(hip):
    c = hip.Context(time_limit_s=120, flags=hip.LOCAL_NET)
    yield c
    c.close()


def evaluate(hip, ctx, code):
    ctx.load(hip.compile_book(code))
    return ctx.eval()


def test_local_reference_programs(golden, hip, lctx):
    for case in golden["programs"]:
        if "compile_error" in case:
            continue
        print("case", case["name"], flush=True)
        r = evaluate(hip, lctx, case["code"])
        assert r.err == 0
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], (case["name"], r.canonical[:200])


def test_local_small_workloads(golden, hip, lctx):
    for case in golden["workloads"]:
        if case.get("full_size"):
            continue
        print("case", case["name"], flush=True)
        code, expect = getattr(P, case["generator"])(**case["params"])
        r = evaluate(hip, lctx, code)
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], case["name"]


@pytest.mark.parametrize("gen", ["sum_tree", "bitonic", "radix", "church"])
def test_local_full_size_workloads(golden, hip, lctx, gen):
    case = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == gen)
    print("case", case["name"], flush=True)
    code, expect = getattr(P, gen)(**case["params"])
    r = evaluate(hip, lctx, code)
    print("leaks", r.leak, "diag", r.diag, flush=True)
    assert r.itrs == case["itrs"]
    assert sha(r.canonical) == case["nf_sha256"]
    if gen == "bitonic":
        assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
    else:
        assert r.result == str(expect)


def test_local_histogram_matches_oracle(hip, lctx, orc):
    for code in [P.sum_tree(10)[0], P.bitonic(6, 2)[0], P.radix(6, 1)[0], P.church(3, 4)[0]]:
        w = hip.compile_book(code)
        o = orc.run_oracle(w)
        lctx.load(w)
        r = lctx.eval()
        assert r.kinds == o.kinds
        assert r.bytes == o.bytes
        assert r.leak > 0   # results reach the HBM root only through LEAKs


@pytest.mark.parametrize("cfg", [dict(turn_steps=1), dict(turn_steps=3, waves=64), dict(waves=1),
                                 dict(idle_pct=101), dict(idle_pct=5),
                                 dict(waves=256, node_cap=1 << 19, vars_cap=1 << 19)])
def test_local_schedule_invariance(golden, hip, cfg):
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    print("cfg", cfg, flush=True)
    r = hip.run_code(code, time_limit_s=120, flags=hip.LOCAL_NET, **cfg)
    assert r.itrs == case["itrs"]
    assert P.flatten_tree(P.parse_tree_text(r.result)) == expect


def test_local_batch_small(hip):
    # several nets booted into one heap, each with its own HBM root var
    code, nets = P.batch(5, 12, chunk=4)
    fids = P.def_fids(code)
    entries = [fids[f"net{i}"] for i in range(len(nets))]
    with hip.Context(time_limit_s=120, flags=hip.LOCAL_NET) as c, hip.Context(time_limit_s=120) as g:
        words = hip.compile_book(code)
        c.load(words)
        g.load(words)
        per_l, tot_l = c.eval_batch(entries, None)
        per_g, tot_g = g.eval_batch(entries, None)
    assert tot_l.itrs == tot_g.itrs
    assert [r.canonical for r in per_l] == [r.canonical for r in per_g]
