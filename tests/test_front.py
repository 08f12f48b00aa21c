"""Front end: parser, builder, serializer and canonical printer.

The C++ front end restates the reference's Rust host (src/ast.rs, src/hvm.rs
to_buffer).  It is pinned two ways:
* end to end: every reference test program and example, compiled by this
  front end and run by the compiled reference run-c, prints exactly the
  normal form in the reference's own insta snapshot (tests/snapshots), which
  the reference produced with its Rust front end (recorded in golden.json);
* by hand-derived book buffers for small programs (fid order, pre-order node
  numbering, first-occurrence var numbering, safety, the `&!` flag).
"""
import pytest

from conftest import ROOT  # noqa: F401


def name_words(name):
    b = name.encode() + b"\0" * (256 - len(name))
    return [int.from_bytes(b[i:i + 4], "little") for i in range(0, 256, 4)]


def test_snapshots_reproduced_through_reference_runtime(golden, hip):
    checked = 0
    for case in golden["programs"]:
        snap = case.get("snapshot")
        if "compile_error" in case:
            assert "missing `@main` definition" in snap
            assert case["compile_error"] == "missing `@main` definition"
            continue
        if snap is None or snap.startswith("ERROR"):
            # safety-check.hvm is a rust-only test: the Rust interpreter refuses to
            # clone an unsafe REF, the C runtime expands it (src/hvm.c:878-880).
            continue
        assert case["nf"] == hip.canonical(snap), case["name"]
        checked += 1
    assert checked >= 12


def test_compile_matches_golden_programs(golden, hip):
    for case in golden["programs"]:
        if "compile_error" in case:
            with pytest.raises(hip.HvmError):
                hip.compile_book(case["code"])
        else:
            w = hip.compile_book(case["code"])
            assert w[0] >= 1


def test_single_def_layout(hip):
    w = hip.compile_book("@main = (a a)")
    # defs_len, fid, name, safe, rbag_len, node_len, vars_len, root, node0
    assert w == [1, 0] + name_words("main") + [1, 0, 1, 1, (0 << 3) | 4, 0, 0]


def test_fid_order_and_safety(hip):
    code = "@main = x & @foo ~ (x *)\n@foo = {a a}\n@Bar = *\n"
    w = hip.compile_book(code)
    assert w[0] == 3
    # main is fid 0; the others follow in byte order: "Bar" < "foo"
    i = 1
    defs = {}
    for _ in range(3):
        fid = w[i]
        name = bytes(b for x in w[i + 1:i + 65] for b in x.to_bytes(4, "little")).rstrip(b"\0").decode()
        safe, rl, nl, vl, root = w[i + 65:i + 70]
        body = w[i + 70:i + 70 + 2 * (rl + nl)]
        defs[fid] = (name, safe, rl, nl, vl, root, body)
        i += 70 + 2 * (rl + nl)
    assert defs[0][0] == "main" and defs[1][0] == "Bar" and defs[2][0] == "foo"
    # foo holds a DUP: unsafe; main references foo: unsafe; Bar: safe
    assert defs[2][1] == 0 and defs[0][1] == 0 and defs[1][1] == 1
    # main: root VAR 0; redex (REF foo = fid 2, CON node 0); node 0 = (VAR 0, ERA)
    name, safe, rl, nl, vl, root, body = defs[0]
    assert (rl, nl, vl, root) == (1, 1, 1, 0)
    assert body == [(2 << 3) | 1, (0 << 3) | 4, 0, 2]


def test_par_flag_and_node_order(hip):
    w = hip.compile_book("@main = (a (b c)) &! @main ~ (c b) & a ~ *")
    # root CON 0 = (VAR a, CON 1); node 1 = (VAR b, VAR c); redex 0: REF main with bit 28
    rl, nl, vl, root = w[67:71]
    assert (rl, nl, vl) == (2, 3, 3)
    assert root == 4
    rbag = w[71:75]
    assert rbag[0] == ((0x10000000 << 3) | 1)
    assert rbag[1] == (2 << 3) | 4          # redex fst: CON node 2 = (c b)
    assert rbag[2:4] == [0 << 3, 2]          # redex 1: VAR a ~ ERA
    nodes = w[75:81]
    assert nodes == [0, (1 << 3) | 4, 1 << 3, 2 << 3, 2 << 3, 1 << 3]


@pytest.mark.parametrize("lit,numb", [
    ("0", (0 << 5) | 1), ("16777215", (16777215 << 5) | 1), ("16777216", 1),   # u24 wraps
    ("+5", (5 << 5) | 2), ("-5", ((-5 & 0xFFFFFFFF) << 5 | 2) & 0xFFFFFFFF),
    ("[+]", 4 << 5), ("[u24]", 1 << 5), ("[*2]", (2 << 5) | 7), ("[:<<]", 0x14 << 5),
    ("1.0", ((0x3F800000 >> 8) << 5) | 3), ("+inf", ((0x7F800000 >> 8) << 5) | 3),
])
def test_number_literals(hip, lit, numb):
    w = hip.compile_book(f"@main = {lit}")
    root = w[70]
    assert root & 7 == 3
    assert root >> 3 == numb & 0x1FFFFFFF


@pytest.mark.parametrize("printed,canon", [
    ("((xa xb) (xb xa))", "((x0 x1) (x1 x0))"),
    ("{2.1500244e+00 -1.2340000e+03}", "{2.1500244 -1234.0}"),
    ("$(8.9087990e-06 1.6775936e+07)", "$(8.908799e-6 16775936.0)"),
    ("?(+NaN -inf)", "?(+NaN -inf)"),
    ("[+0x000000A]", "[+0x000000A]"),
    ("(@foo *)", "(@foo *)"),
])
def test_canonical(hip, printed, canon):
    assert hip.canonical(printed) == canon


def test_parse_errors(hip):
    for bad in ["@main = (a", "main = *", "@main = $(a b", "@main = x & @nope ~ x"]:
        with pytest.raises(hip.HvmError):
            hip.compile_book(bad)


def test_io_fixture_programs_compile(hip):
    """tests/golden/io.json (the reference run-c's output on IO programs):
    every program compiles, and the fixture records a result and ITRS for
    each, so the GPU IO test compares against real reference output."""
    import json
    import os
    with open(os.path.join(ROOT, "tests", "golden", "io.json")) as f:
        fx = json.load(f)
    names = [c["name"] for c in fx["programs"]]
    assert "tests/programs/io/basic.hvm" in names and "tests/programs/io/echo_seek.hvm" in names
    for c in fx["programs"]:
        assert hip.compile_book(c["code"])
        assert c["result"] and c["itrs"] > 0 and c["returncode"] == 0
    # The program of this repository is the one the fixture was made from.
    with open(os.path.join(ROOT, "tests", "programs", "io", "echo_seek.hvm")) as f:
        assert f.read() == next(c for c in fx["programs"] if c["name"].endswith("echo_seek.hvm"))["code"]
    # Its stdin READ and the 5 bytes read at offset 33 of the LICENSE fixture.
    echo = next(c for c in fx["programs"] if c["name"].endswith("echo_seek.hvm"))
    assert echo["lines"] == [fx["stdin"][:12] + fx["files"]["LICENSE"][33:38]]


def test_io_fixture_matches_reference_snapshots():
    """run-c's output on the reference's IO programs (tests/golden/io.json)
    equals the reference's own snapshots of them (tests/run.rs:67-110 runs
    them from the repository root, whose LICENSE the fixture starts like)."""
    import json
    import os
    with open(os.path.join(ROOT, "tests", "golden", "io.json")) as f:
        fx = json.load(f)
    snaps = [c for c in fx["programs"] if "reference_snapshot" in c]
    assert len(snaps) == 6
    for c in snaps:
        assert c["lines"] == c["reference_snapshot"]["lines"], c["name"]
        assert c["result"] == c["reference_snapshot"]["result"], c["name"]
