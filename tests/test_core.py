"""Value encodings, rule table, f24 packing and the integer ALU (hvm_core.h).

Pinned against the reference's literal tables and known answers:
* rule table: the 8x8 table of src/hvm.c:319-333 (also src/hvm.rs:148-158);
* priorities: src/hvm.c:345-348;
* f24 literal packing: every vector of `test_f24` (src/hvm.rs:1160-1207);
* f24 runtime packing: the C/CUDA runtimes' rounding term uses a logical not
  (src/hvm.c:402, src/hvm.cu:508), so an exact tie rounds up unless the kept
  bits are zero; checked on tie cases where the two rules differ;
* integer ALU: closed-form u24/i24 semantics, plus the answers commented in
  tests/programs/numerics/u24.hvm and i24.hvm (10 op 2).
"""
import json
import os
import subprocess

import pytest

from conftest import ROOT

LINK, CALL, VOID, ERAS, ANNI, COMM, OPER, SWIT = range(8)
# src/hvm.c:319-333, rows = tag(a), columns = tag(b): VAR REF ERA NUM CON DUP OPR SWI
REF_TABLE = [
    [LINK, LINK, LINK, LINK, LINK, LINK, LINK, LINK],
    [LINK, VOID, VOID, VOID, CALL, CALL, CALL, CALL],
    [LINK, VOID, VOID, VOID, ERAS, ERAS, ERAS, ERAS],
    [LINK, VOID, VOID, VOID, ERAS, ERAS, OPER, SWIT],
    [LINK, CALL, ERAS, ERAS, ANNI, COMM, COMM, COMM],
    [LINK, CALL, ERAS, ERAS, COMM, ANNI, COMM, COMM],
    [LINK, CALL, ERAS, OPER, COMM, COMM, ANNI, COMM],
    [LINK, CALL, ERAS, SWIT, COMM, COMM, COMM, ANNI],
]


# (input bits, expected bits) pairs from test_f24 (src/hvm.rs:1182-1198)
TIE_VECTORS = [
    (0b00_00000000, 0b00_00000000), (0b00_01111111, 0b00_00000000), (0b00_10000000, 0b00_00000000),
    (0b00_10000001, 0b01_00000000), (0b00_11111111, 0b01_00000000), (0b01_00000000, 0b01_00000000),
    (0b01_01111111, 0b01_00000000), (0b01_10000000, 0b10_00000000), (0b01_10000001, 0b10_00000000),
    (0b01_11111111, 0b10_00000000),
]
# round-trip values of test_f24 (src/hvm.rs:1167-1181), as f32 bits
ROUND_TRIP = [0x00000000, 0x80000000, 0x3F800000, 0xBF800000, 0x3F8CCD00, 0xBF8CCD00,
              0x7F800000, 0xFF800000, 0x00000100, 0x80000100, 0x00007B00, 0x80007B00,
              0x007FFF00, 0x00800000]


@pytest.fixture(scope="module")
def core(tmp_path_factory):
    d = tmp_path_factory.mktemp("core")
    exe = str(d / "core_check")
    subprocess.run(["gcc", "-O1", "-std=gnu11", os.path.join(ROOT, "tests", "core_check.c"), "-o", exe, "-lm"],
                   check=True)
    args = [f"{b:x}" for b, _ in TIE_VECTORS] + [f"{b:x}" for b in ROUND_TRIP]
    args += ["7fc00000", "7f800001", "ff800001", "7fffffff", "3f800280", "3f800180"]
    out = subprocess.run([exe] + args, check=True, capture_output=True, text=True).stdout
    return json.loads(out), args


def test_rule_table(core):
    data, _ = core
    got = [data["rules"][a * 8: a * 8 + 8] for a in range(8)]
    assert got == REF_TABLE


def test_priorities(core):
    data, _ = core
    # LINK, VOID, ERAS, ANNI are high priority (0b00011101)
    assert data["high"] == [1, 0, 1, 1, 1, 0, 0, 0]


def test_f24_literal_matches_rust_test_vectors(core):
    data, args = core
    lit = data["f24_literal"]
    for i, (_, exp) in enumerate(TIE_VECTORS):
        assert lit[i] == exp, (i, hex(lit[i]), hex(exp))
    base = len(TIE_VECTORS)
    for j, b in enumerate(ROUND_TRIP):
        assert lit[base + j] == b
    nan_base = base + len(ROUND_TRIP)
    for k in range(4):  # NaN inputs stay NaN
        v = lit[nan_base + k]
        assert (v >> 23) & 0xFF == 0xFF and v & 0x7FFFFF != 0


def test_f24_runtime_rounding_differs_only_on_ties(core):
    data, args = core
    lit, rt = data["f24_literal"], data["f24_runtime"]
    # 0x3f800280: kept bits odd-free? low byte 0x80 is an exact tie, kept bits 0x3f8002 (even, nonzero):
    # ties-to-even keeps 0x3f8002, the runtime rounds up to 0x3f8003.
    i = args.index("3f800280")
    assert lit[i] == 0x3F800200 and rt[i] == 0x3F800300
    # kept bits odd: both round up
    j = args.index("3f800180")
    assert lit[j] == rt[j] == 0x3F800200
    # non-ties agree
    for k, (a, b) in enumerate(zip(lit, rt)):
        if k not in (i,):
            bits = int(args[k], 16)
            if (bits & 0xFF) != 0x80:
                assert a == b, args[k]


# A Numb lives in the 29-bit value of a NUM port (new_port truncates val << 3),
# so ALU results are compared modulo 2^29.
P29 = (1 << 29) - 1


def u24(x):
    return ((x << 5) | 1) & P29


def i24(x):
    return (((x & 0xFFFFFFFF) << 5) | 2) & P29


def test_integer_alu(core):
    data, _ = core
    ops = data["ops"]
    a, b = 10, 2
    expect_u = {
        4: a + b, 5: a - b, 7: a * b, 8: a // b, 10: a % b, 12: int(a == b), 13: int(a != b),
        14: int(a < b), 15: int(a > b), 16: a & b, 17: a | b, 18: a ^ b, 19: a << b, 21: a >> b,
        6: (b - a) % (1 << 24), 9: b // a, 11: b % a, 20: b << a, 22: b >> a,
    }
    for op, v in expect_u.items():
        assert ops[str(op)][0] & P29 == u24(v % (1 << 24)), op
    # answers commented in tests/programs/numerics/u24.hvm (10 op 2)
    commented = {4: 12, 5: 8, 7: 20, 8: 5, 10: 0, 12: 0, 13: 1, 14: 0, 15: 1, 16: 2, 17: 10, 18: 8, 19: 40, 21: 2}
    for op, v in commented.items():
        assert ops[str(op)][0] & P29 == u24(v)
    # i24: -10 op 3, truncating division like C
    x, y = -10, 3
    expect_i = {4: i24(x + y), 5: i24(x - y), 6: i24(y - x), 7: i24(x * y), 8: i24(-3), 9: i24(0),
                10: i24(-1), 11: i24(3), 12: u24(0), 13: u24(1), 14: u24(1), 15: u24(0),
                16: i24(x & y), 17: i24(x | y), 18: i24(x ^ y), 19: i24(0), 20: i24(0), 21: i24(0), 22: i24(0)}
    for op, v in expect_i.items():
        assert ops[str(op)][1] & P29 == v, op
