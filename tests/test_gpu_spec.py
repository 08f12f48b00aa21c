"""Compiled CALLs (HVM_HIP_COMPILED_CALLS) on the GPU, against run-c.

With the flag the evaluator loads every def's template pre-reduced (REF~node
CALLs inlined, same-tag ANNIs performed) and adds each def's static CALL and
ANNI counts at every expansion.  Bar: the same as the plain evaluator — ITRS
and the canonical normal form bit-identical to the compiled reference run-c
(tests/golden/golden.json), the per-kind histogram and canonical bytes equal
to the oracle's — on every reference program and on the workloads, under
schedule variations and with the local-net evaluator.
"""
import hashlib

import pytest

import programs as P

pytestmark = pytest.mark.gpu


def sha(s):
    return hashlib.sha256(s.encode()).hexdigest()


@pytest.fixture(scope="module")
def cctx(hip):
    c = hip.Context(time_limit_s=120, flags=hip.COMPILED_CALLS)
    yield c
    c.close()


def run(hip, ctx, code):
    ctx.load(hip.compile_book(code))
    return ctx.eval()


def test_reference_programs(golden, hip, cctx):
    for case in golden["programs"]:
        if "compile_error" in case:
            continue
        print("case", case["name"], flush=True)
        r = run(hip, cctx, case["code"])
        assert r.err == 0
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], (case["name"], r.canonical[:200])


def test_small_workloads(golden, hip, cctx):
    for case in golden["workloads"]:
        if case.get("full_size"):
            continue
        print("case", case["name"], flush=True)
        code, _ = getattr(P, case["generator"])(**case["params"])
        r = run(hip, cctx, code)
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], case["name"]


@pytest.mark.parametrize("gen", ["sum_tree", "bitonic", "radix", "church"])
def test_full_size_workloads(golden, hip, cctx, gen):
    case = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == gen)
    print("case", case["name"], flush=True)
    code, expect = getattr(P, gen)(**case["params"])
    r = run(hip, cctx, code)
    assert r.itrs == case["itrs"]
    assert sha(r.canonical) == case["nf_sha256"]
    if gen == "bitonic":
        assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
    else:
        assert r.result == str(expect)


def test_histogram_and_bytes_match_oracle(hip, cctx, orc):
    for code in [P.sum_tree(10)[0], P.bitonic(6, 2)[0], P.radix(6, 1)[0], P.church(3, 4)[0]]:
        w = hip.compile_book(code)
        o = orc.run_oracle(w)
        cctx.load(w)
        r = cctx.eval()
        assert r.kinds == o.kinds
        assert r.bytes == o.bytes


@pytest.mark.parametrize("cfg", [dict(waves=1), dict(turn_steps=1), dict(turn_steps=3, waves=64),
                                 dict(waves=256, node_cap=1 << 19, vars_cap=1 << 19),
                                 dict(flags=8)])
def test_schedule_invariance(golden, hip, cfg):
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    cfg = dict(cfg, flags=cfg.get("flags", 0) | hip.COMPILED_CALLS)
    r = hip.run_code(code, time_limit_s=120, **cfg)
    assert r.itrs == case["itrs"]
    assert P.flatten_tree(P.parse_tree_text(r.result)) == expect


def test_same_result_as_plain_evaluator(hip, cctx):
    code = P.bitonic(10, 0)[0]
    w = hip.compile_book(code)
    with hip.Context(time_limit_s=120) as plain:
        plain.load(w)
        a = plain.eval()
    cctx.load(w)
    b = cctx.eval()
    assert (a.itrs, a.kinds, a.bytes, a.canonical) == (b.itrs, b.kinds, b.bytes, b.canonical)


def stress_code(depth, length):
    """examples/stress/main.hvm of the reference with its sizes as parameters:
    2^depth sequential loops of `length` iterations, summed."""
    return f"""@fun = (?((@fun__C0 @fun__C1) a) a)
@fun__C0 = a
  & @loop ~ ({length} a)
@fun__C1 = ({{a b}} d)
  &! @fun ~ (a $([+] $(c d)))
  &! @fun ~ (b c)
@loop = (?((0 @loop__C0) a) a)
@loop__C0 = a
  & @loop ~ a
@main = a
  & @fun ~ ({depth} a)
"""


@pytest.mark.parametrize("depth,length", [(4, 300), (10, 64)])
def test_fast_calls_on_loops(hip, orc, depth, length):
    # Every loop iteration is a CALL whose caller passes (n, k): the fast
    # CALL performs the ANNI, the SWITCH and the branch ANNI in the CALL.
    # Same counts of every kind and the same bytes as the oracle's.
    w = hip.compile_book(stress_code(depth, length))
    o = orc.run_oracle(w)
    with hip.Context(time_limit_s=120) as plain:
        plain.load(w)
        a = plain.eval()
    with hip.Context(time_limit_s=120, flags=hip.COMPILED_CALLS) as comp:
        comp.load(w)
        b = comp.eval()
    assert a.kinds == b.kinds == o.kinds
    assert a.bytes == b.bytes == o.bytes
    assert a.canonical == b.canonical == hip.canonical(o.result)
    assert b.diag["steps"] < a.diag["steps"]
