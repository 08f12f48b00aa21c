"""The 64-bit-port evaluator (HVM_HIP_WIDE) on the GPU, against run-c.

Ports of 64 bits and nodes of two 64-bit halves lift the reference's 2^29
slot cap (paper/HVM2.typst:1212-1222).  Numbers keep the 29-bit encoding, so
the bar is the 32-bit evaluator's: ITRS and the canonical normal form
bit-identical to the compiled reference run-c (tests/golden/golden.json), the
per-kind histogram and canonical bytes equal to the oracle's.  Beyond parity:
nets whose slots lie above 2^29 (the waves' slices of a 2^31-slot net start
past it), the largest capacity (2^32 - 2 slots), the shared pool, one-step
turns, and the error paths.
"""
import hashlib

import pytest

import programs as P

pytestmark = pytest.mark.gpu


def sha(s):
    return hashlib.sha256(s.encode()).hexdigest()


@pytest.fixture(scope="module")
def wctx(hip):
    c = hip.Context(time_limit_s=120, flags=hip.WIDE, node_cap=1 << 26, vars_cap=1 << 26)
    yield c
    c.close()


def run(hip, ctx, code):
    ctx.load(hip.compile_book(code))
    return ctx.eval()


def test_reference_programs(golden, hip, wctx):
    for case in golden["programs"]:
        if "compile_error" in case:
            continue
        print("case", case["name"], flush=True)
        r = run(hip, wctx, case["code"])
        assert r.err == 0
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], (case["name"], r.canonical[:200])


def test_small_workloads(golden, hip, wctx):
    for case in golden["workloads"]:
        if case.get("full_size"):
            continue
        print("case", case["name"], flush=True)
        code, _ = getattr(P, case["generator"])(**case["params"])
        r = run(hip, wctx, code)
        assert r.itrs == case["itrs"], case["name"]
        assert sha(r.canonical) == case["nf_sha256"], case["name"]


@pytest.mark.parametrize("gen", ["sum_tree", "bitonic", "radix", "church"])
def test_full_size_workloads(golden, hip, gen):
    case = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == gen)
    print("case", case["name"], flush=True)
    code, expect = getattr(P, gen)(**case["params"])
    r = hip.run_code(code, time_limit_s=120, flags=hip.WIDE)
    assert r.itrs == case["itrs"]
    assert sha(r.canonical) == case["nf_sha256"]
    if gen == "bitonic":
        assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
    else:
        assert r.result == str(expect)


def test_histogram_and_bytes_match_oracle(hip, wctx, orc):
    for code in [P.sum_tree(10)[0], P.bitonic(6, 2)[0], P.radix(6, 1)[0], P.church(3, 4)[0]]:
        w = hip.compile_book(code)
        o = orc.run_oracle(w)
        wctx.load(w)
        r = wctx.eval()
        assert r.kinds == o.kinds
        assert r.bytes == o.bytes


@pytest.mark.parametrize("cfg", [dict(waves=1), dict(turn_steps=1), dict(turn_steps=3, waves=64),
                                 # tiny slices: allocation overflows into the shared pool
                                 dict(waves=64, node_cap=1 << 16, vars_cap=1 << 16),
                                 dict(idle_pct=101), dict(flags=16)])
def test_schedule_invariance(golden, hip, cfg):
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    cfg = dict(cfg, flags=cfg.get("flags", 0) | hip.WIDE)
    r = hip.run_code(code, time_limit_s=120, **cfg)
    assert r.itrs == case["itrs"]
    assert P.flatten_tree(P.parse_tree_text(r.result)) == expect


def test_slots_above_2_29(golden, hip):
    # A 2^31-slot net over 8 waves: every wave but the first three allocates
    # only slots past 2^29, which 32-bit ports cannot address.
    case = next(c for c in golden["workloads"] if c.get("full_size") and c["generator"] == "bitonic")
    code, expect = P.bitonic(16, 0)
    r = hip.run_code(code, time_limit_s=120, flags=hip.WIDE, waves=8, node_cap=1 << 31, vars_cap=1 << 31)
    assert r.itrs == case["itrs"]
    assert sha(r.canonical) == case["nf_sha256"]


def test_largest_capacity(golden, hip):
    case = next(c for c in golden["workloads"] if c["name"] == "bitonic_k8_seed3")
    code, expect = P.bitonic(8, 3)
    cap = (1 << 32) - 2
    r = hip.run_code(code, time_limit_s=120, flags=hip.WIDE, node_cap=cap, vars_cap=cap)
    assert r.itrs == case["itrs"]
    assert P.flatten_tree(P.parse_tree_text(r.result)) == expect


def test_capacity_limits_rejected(hip):
    with pytest.raises(hip.HvmError):
        hip.Context(flags=hip.WIDE, node_cap=1 << 32)
    with pytest.raises(hip.HvmError):
        hip.Context(flags=hip.WIDE | hip.LOCAL_NET)


def test_out_of_memory_is_reported(hip):
    code, _ = P.bitonic(12, 0)
    with pytest.raises(hip.HvmError) as e:
        hip.run_code(code, time_limit_s=60, flags=hip.WIDE, waves=16, node_cap=1 << 14, vars_cap=1 << 14)
    import re
    m = re.search(r"(?:flags|err=)\s*0x([0-9a-f]+)", str(e.value))
    assert m and int(m.group(1), 16) & 3, str(e.value)


def test_net_beyond_2_29_live_nodes(hip):
    # tree_gen(30): the normal form alone is 2^30 - 1 CON nodes.  The 32-bit
    # evaluator's largest net (2^29 slots) runs out of nodes; the wide one
    # finishes with the interaction counts the closed forms give (pinned by
    # the oracle and run-c at small depths).
    d = 30
    code, _ = P.tree_gen(d)
    w = hip.compile_book(code)
    with pytest.raises(hip.HvmError):
        hip.run_hip(w, time_limit_s=100, flags=hip.NO_READBACK, node_cap=1 << 29, vars_cap=1 << 29)
    # up to ~10^8 redexes wait between turns: 2^28-entry bags (4 GiB each)
    cap = (1 << 32) - 2
    r = hip.run_hip(w, time_limit_s=100, flags=hip.WIDE | hip.NO_READBACK, node_cap=cap, vars_cap=cap,
                    bag_cap=1 << 28)
    assert r.kinds == P.tree_gen_kinds(d)
    assert r.itrs == 13 * 2 ** d - 7
