"""Parity of the HIP evaluator with the reference (run-c), on the GPU.

Bar: the canonical normal form (tests/run.rs:112-146 renaming) and the
interaction count ITRS are bit-identical to the compiled reference run-c,
recorded in tests/golden/golden.json by tests/golden/make_golden.py.  Full
sizes are additionally checked against closed forms (2^20, the sorted keys,
Σ sorted[i]·i mod 2^24, 2^24 mod 2^24).  Every test goes through the C ABI
(libhvm_hip.so); there is no CPU fallback behind it.
"""
import hashlib

import pytest

import programs as P

pytestmark = pytest.mark.gpu

KINDS_BYTES = {"CALL_COPY": 32, "VOID": 0, "ERAS": 32, "ANNI": 48, "COMM": 112,
               "OPER1": 32, "OPER2": 40, "SWIT0": 32, "SWITN": 40}


def sha(s):
    return hashlib.sha256(s.encode()).hexdigest()


@pytest.fixture(scope="module")
def ctx(hip):
    c = hip.Context(time_limit_s=120)
    yield c
    c.close()


def evaluate(hip, ctx, code):
    ctx.load(hip.compile_book(code))
    return ctx.eval()


def test_reference_programs(golden, hip, ctx):
    for case in golden["programs"]:
        if "compile_error" in case:
            continue
        print("case", case["name"], flush=True)
        r = evaluate(hip, ctx, case["code"])
        assert r.err == 0
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], (case["name"], r.canonical[:200])
        assert sum(r.kinds.values()) == r.itrs


def test_small_workloads(golden, hip, ctx):
    for case in golden["workloads"]:
        if case.get("full_size"):
            continue
        print("case", case["name"], flush=True)
        code, expect = getattr(P, case["generator"])(**case["params"])
        r = evaluate(hip, ctx, code)
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], case["name"]
        if case["generator"] == "bitonic":
            assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
        else:
            assert r.result == str(expect)


# stress: the reference's stress program with 2^18 loops of 1024 iterations,
# whose frontier (19 M redexes) overflowed the fixed 2^24-entry bags; bags
# are now sized from the node capacity
@pytest.mark.parametrize("gen", ["sum_tree", "bitonic", "radix", "church", "radix_lsd", "church_loop", "stress"])
def test_full_size_workloads(golden, hip, ctx, gen):
    case = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == gen)
    print("case", case["name"], flush=True)
    code, expect = getattr(P, gen)(**case["params"])
    r = evaluate(hip, ctx, code)
    assert r.itrs == case["itrs"]
    assert sha(r.canonical) == case["nf_sha256"]
    if gen == "bitonic":
        got = P.flatten_tree(P.parse_tree_text(r.result))
        assert got == expect                     # sorted, key for key
        assert all(a <= b for a, b in zip(got, got[1:]))
    else:
        assert r.result == str(expect)


def test_histogram_matches_oracle(hip, ctx, orc):
    for code in [P.sum_tree(10)[0], P.bitonic(6, 2)[0], P.radix(6, 1)[0], P.church(3, 4)[0],
                 P.radix_lsd(6, 3)[0], P.church_loop(4, 6)[0]]:
        w = hip.compile_book(code)
        o = orc.run_oracle(w)
        ctx.load(w)
        r = ctx.eval()
        assert r.kinds == o.kinds
        assert r.bytes == o.bytes


@pytest.mark.parametrize("cfg", [dict(waves=1), dict(waves=7), dict(turn_steps=1),
                                 dict(turn_steps=3, waves=64), dict(node_cap=1 << 22, vars_cap=1 << 22),
                                 dict(idle_pct=101), dict(idle_pct=5), dict(flags=2),
                                 # small slices: waves overflow into the shared pool
                                 dict(waves=256, node_cap=1 << 19, vars_cap=1 << 19)])
def test_schedule_invariance(golden, hip, cfg):
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    r = hip.run_code(code, time_limit_s=120, **cfg)
    assert r.itrs == case["itrs"]
    assert P.flatten_tree(P.parse_tree_text(r.result)) == expect


def test_context_reuse_is_repeatable(golden, hip, ctx):
    code, _ = P.radix(8, 2)
    ctx.load(hip.compile_book(code))
    a = ctx.eval()
    b = ctx.eval()
    assert (a.itrs, a.canonical) == (b.itrs, b.canonical)


def test_oversized_def_is_rejected(hip, ctx):
    leaves = " ".join("*" for _ in range(600))
    tree = "*"
    for _ in range(600):
        tree = f"(* {tree})"
    with pytest.raises(hip.HvmError):
        ctx.load(hip.compile_book(f"@main = {tree}"))


# Failure handling, forced from the host: a net or bag far too small for the
# program.  The evaluator must stop with the matching error flag (every kernel
# exits normally) and report it through the status code, never a result.
OOM_NODE, OOM_VARS, BAG_OVERFLOW = 1, 2, 4


def error_flags(e) -> int:
    """The evaluator's error flags named in an HvmError message."""
    import re
    m = re.search(r"(?:flags|err=)\s*0x([0-9a-f]+)", str(e.value))
    assert m, str(e.value)
    return int(m.group(1), 16)


def test_out_of_memory_is_reported(hip):
    code, _ = P.bitonic(12, 0)   # peak live nodes far above the 2^15-slot net
    with pytest.raises(hip.HvmError) as e:
        hip.run_code(code, time_limit_s=60, waves=64, node_cap=1 << 15, vars_cap=1 << 15)
    assert error_flags(e) & (OOM_NODE | OOM_VARS)


def test_bag_overflow_is_reported(hip):
    code, _ = P.sum_tree(16)
    with pytest.raises(hip.HvmError) as e:
        hip.run_code(code, time_limit_s=60, bag_cap=64)
    assert error_flags(e) & BAG_OVERFLOW


def test_context_recovers_after_an_error(golden, hip):
    # an error leaves the context usable: the next evaluation starts clean
    case = next(c for c in golden["workloads"] if c["generator"] == "sum_tree" and c["params"] == {"depth": 4})
    with hip.Context(time_limit_s=60, bag_cap=64) as ctx:
        ctx.load(hip.compile_book(P.sum_tree(16)[0]))
        with pytest.raises(hip.HvmError):
            ctx.eval()
        ctx.load(hip.compile_book(P.sum_tree(4)[0]))
        r = ctx.eval()
        assert r.err == 0 and r.itrs == case["itrs"] and r.result == "16"


def test_graph_replay_matches_direct_launches(golden, hip):
    # HVM_HIP_GRAPHS: chunks of turns replayed from captured graphs, many
    # replays per evaluation, several evaluations (with eager host work and
    # synchronisations between them) on one context; every one identical to
    # the run-c record and to direct launches.
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    w = hip.compile_book(code)
    for flags in (hip.GRAPHS, hip.GRAPHS | hip.COMPILED_CALLS):
        with hip.Context(time_limit_s=120, flags=flags, turn_steps=7) as ctx:
            ctx.load(w)
            for _ in range(3):
                r = ctx.eval()
                assert r.itrs == case["itrs"]
                assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
    full = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == "radix")
    code, expect = P.radix(**full["params"])
    r = hip.run_code(code, time_limit_s=120, flags=hip.GRAPHS)
    assert r.itrs == full["itrs"] and r.result == str(expect)
