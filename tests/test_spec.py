"""Compiled CALLs: the load-time specialisation of def templates, on the CPU.

The specialised book is run by the oracle (the reference's interpreter
restated) with its expansions counted per def.  Strong confluence of
interaction nets fixes what must hold, and a wrong inlining, a dropped or
double-counted static interaction, or a mis-linked var fails one of:
  * the normal form equals the original book's (canonical names);
  * ITRS of the original = ITRS of the specialised book + Σ over defs of
    (its expansions × its static CALLs + ANNIs);
  * the same per kind: CALL and ANNI gain exactly the static counts, every
    other kind is unchanged;
  * the canonical bytes likewise.
The GPU evaluator adds the same static counts per expansion
(tests/test_gpu_spec.py checks it against run-c's ITRS).
"""
import pytest

import programs as P


def cases(golden):
    out = []
    for c in golden["programs"]:
        if "compile_error" in c or c.get("itrs") is None:
            continue
        out.append((c["name"], c["code"]))
    out += [("bitonic(8)", P.bitonic(8, 0)[0]), ("radix(8)", P.radix(8, 0)[0]),
            ("sum_tree(10)", P.sum_tree(10)[0]), ("church(5,5)", P.church(5, 5)[0])]
    return out


def own_bytes(words):
    """8N + 4V + 8(R+1) of every def of a wire buffer (src/hvm.rs:957-1003:
    fid, 64-word name, safe, rlen, nlen, vlen, root, rbag pairs, node pairs)."""
    out, i = [], 1
    for _ in range(words[0]):
        rlen, nlen, vlen = words[i + 66], words[i + 67], words[i + 68]
        out.append(8 * nlen + 4 * vlen + 8 * (rlen + 1))
        i += 70 + 2 * (rlen + nlen)
    return out


def check(hip, orc, code, budget):
    words = hip.compile_book(code)
    base = orc.run_oracle(words)
    assert base.status == 0
    sw, counts = hip.specialize(words, budget)
    r = orc.run_oracle(sw, calls=True)
    assert r.status == 0
    assert hip.canonical(r.result) == hip.canonical(base.result)
    x_call = sum(n * c[0] for n, c in zip(r.calls, counts))
    x_anni = sum(n * c[1] for n, c in zip(r.calls, counts))
    x_bytes = sum(n * (c[2] - own) for n, c, own in zip(r.calls, counts, own_bytes(sw)) if c[0] or c[1])
    assert r.kinds["CALL"] + x_call == base.kinds["CALL"]
    assert r.kinds["ANNI"] + x_anni == base.kinds["ANNI"]
    for k in base.kinds:
        if k not in ("CALL", "ANNI"):
            assert r.kinds[k] == base.kinds[k], k
    assert r.itrs + x_call + x_anni == base.itrs
    assert r.bytes + x_bytes == base.bytes
    return base, r, counts


def test_reference_programs_and_workloads(golden, hip, orc):
    folded = 0
    for name, code in cases(golden):
        base, r, counts = check(hip, orc, code, hip.SPEC_NODES)
        folded += base.itrs - r.itrs
    assert folded > 0


@pytest.mark.parametrize("budget", [0, 8, 256])
def test_budgets(golden, hip, orc, budget):
    for name, code in cases(golden)[-4:]:
        check(hip, orc, code, budget)


def test_bitonic_folds_annihilations(hip, orc):
    # The argument spines of the sort's calls meet the callee's lambda spines
    # inside the caller's template: those ANNIs become static (the ones after
    # a SWITCH picks a branch stay dynamic).
    base, r, counts = check(hip, orc, P.bitonic(8, 0)[0], hip.SPEC_NODES)
    assert r.kinds["ANNI"] < 0.7 * base.kinds["ANNI"]
    assert r.kinds["CALL"] < base.kinds["CALL"]


def test_recursive_def_terminates(hip, orc):
    code = "@main = r & @f ~ (3 r)\n@f = (?((7 @s) r) r)\n@s = (p r) & @f ~ (p r)\n"
    base, r, counts = check(hip, orc, code, 256)
    assert hip.canonical(base.result) == "7"


def test_malformed_budget_rejected(hip):
    with pytest.raises(hip.HvmError):
        hip.specialize(hip.compile_book("@main = 1"), 257)
