"""The CPU oracle (oracle/hvm_oracle.c) against the reference.

* every reference program: normal form and ITRS equal to the compiled
  reference run-c (golden.json), which in turn equals the reference's insta
  snapshots (test_front.py);
* synthetic workloads at sizes the oracle runs in seconds: same, plus their
  closed forms (2^d, sorted keys, Σ sorted[i]·i, 2^(n+m));
* the per-kind histogram sums to ITRS, and its byte total follows the
  per-kind byte table;
* when oracle/_ref is present (build container), the reference run-c itself is
  checked to be deterministic in ITRS across thread counts.
"""
import hashlib

import pytest

import programs as P

KIND_BYTES = {"CALL_COPY": 32, "VOID": 0, "ERAS": 32, "ANNI": 48, "COMM": 112,
              "OPER1": 32, "OPER2": 40, "SWIT0": 32, "SWITN": 40}


def run(hip, orc, code):
    return orc.run_oracle(hip.compile_book(code))


def test_reference_programs(golden, hip, orc):
    n = 0
    for case in golden["programs"]:
        if "compile_error" in case:
            continue
        o = run(hip, orc, case["code"])
        assert o.status == 0
        assert o.itrs == case["itrs"], case["name"]
        canon = hip.canonical(o.result)
        assert hashlib.sha256(canon.encode()).hexdigest() == case["nf_sha256"], case["name"]
        if "nf" in case:
            assert canon == case["nf"], case["name"]
        assert sum(o.kinds.values()) == o.itrs
        n += 1
    assert n >= 13


def small_workloads(golden, limit_itrs=5_000_000):
    return [c for c in golden["workloads"] if c["itrs"] <= limit_itrs]


def test_small_workloads(golden, hip, orc):
    import hashlib
    cases = small_workloads(golden)
    assert len(cases) >= 12
    for case in cases:
        code, expect = getattr(P, case["generator"])(**case["params"])
        o = run(hip, orc, code)
        assert o.status == 0
        assert o.itrs == case["itrs"], case["name"]
        canon = hip.canonical(o.result)
        assert hashlib.sha256(canon.encode()).hexdigest() == case["nf_sha256"], case["name"]
        if case["generator"] == "bitonic":
            assert P.flatten_tree(P.parse_tree_text(o.result)) == expect
        else:
            assert o.result == str(expect)


def test_histogram_bytes(golden, hip, orc):
    code, _ = P.sum_tree(6)
    o = run(hip, orc, code)
    # sum(d): one CALL @main, 2^(d+1)-1 CALL @sum, 2^d - 1 CALL @sum_s
    d = 6
    assert o.kinds["SWIT0"] == 2 ** d
    assert o.kinds["SWITN"] == 2 ** d - 1
    assert o.kinds["CALL"] == 1 + (2 ** (d + 1) - 1) + (2 ** d - 1)
    # each of the 2^d - 1 additions r1 ~ $([+] $(c r)): OPER-1 makes [+r1], OPER-2
    # parks it on the unresolved c, OPER-1 again when c arrives
    assert o.kinds["OPER1"] == 2 * (2 ** d - 1)
    assert o.kinds["OPER2"] == 2 ** d - 1
    # each recursive step duplicates n-1 once (NUM ~ DUP erases into two copies)
    assert o.kinds["ERAS"] == 2 ** d - 1
    fixed = sum(o.kinds[k] * b for k, b in KIND_BYTES.items())
    assert o.bytes > fixed  # CALL bytes come on top


def test_radix_closed_form_catches_a_wrong_key(hip, orc):
    # the checksum pins the whole sorted order: perturbing one key changes it
    keys = P.keys_distinct(16, 7)
    good = P.radix_checksum(keys)
    bad = P.radix_checksum(keys[:-1] + [keys[-1] ^ 1])
    assert good != bad


@pytest.mark.skipif(not __import__("oracle").have_reference(1), reason="oracle/_ref not built here")
def test_reference_itrs_is_schedule_invariant(hip, orc):
    code, _ = P.bitonic(8, 3)
    w = hip.compile_book(code)
    counts = {orc.run_reference(w, t).itrs for t in (1, 2, 4, 8)}
    assert len(counts) == 1


def test_tree_gen_closed_forms(hip, orc):
    """tree_gen's per-kind counts and normal form as closed forms of the depth,
    pinned against the oracle at depths 0-12 and against the reference run-c
    (where built) at depths 4 and 10: the GPU capacity test uses them at
    depth 30, past what a 2^29-slot net holds."""
    import programs as P
    for d in range(13):
        code, nf = P.tree_gen(d)
        o = orc.run_oracle(hip.compile_book(code))
        assert o.status == 0
        assert o.kinds == P.tree_gen_kinds(d), d
        assert o.itrs == 13 * 2 ** d - 7
        assert o.result == nf
    if orc.have_reference(1):
        for d in (4, 10):
            ref = orc.run_reference(hip.compile_book(P.tree_gen(d)[0]), threads=1)
            assert ref.itrs == 13 * 2 ** d - 7
