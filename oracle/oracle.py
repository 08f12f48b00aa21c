"""CPU oracle bindings.  TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker of the HIP evaluator; never by the product.

* `run_oracle(words)` — the C restatement in hvm_oracle.c (single thread,
  follows src/hvm.c + src/run.c; see hvm_oracle.h for the line map).
* `run_reference(words, threads)` — the reference C runtime itself, compiled
  from /root/reference/src/run.c into _ref/hvm_ref_t<threads> by this
  directory's Makefile, run as a child process; its stdout is parsed the way
  the reference's test harness does (tests/run.rs:112-133).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile
from dataclasses import dataclass, field
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhvm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

KINDS = ["CALL", "CALL_COPY", "VOID", "ERAS", "ANNI", "COMM", "OPER1", "OPER2", "SWIT0", "SWITN"]


class _Stats(C.Structure):
    _fields_ = [("itrs", C.c_uint64), ("kind", C.c_uint64 * 10), ("bytes", C.c_uint64),
                ("normalizes", C.c_uint64), ("peak_nodes", C.c_uint64), ("seconds", C.c_double)]


@dataclass
class OracleResult:
    status: int
    result: Optional[str] = None
    itrs: int = 0
    kinds: dict = field(default_factory=dict)
    bytes: int = 0
    normalizes: int = 0
    peak_nodes: int = 0
    seconds: float = 0.0
    calls: Optional[list] = None   # expansions per fid (run_oracle(..., calls=True))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        L.hvm_oracle_run.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_void_p),
                                     C.POINTER(_Stats)]
        L.hvm_oracle_run.restype = C.c_int
        L.hvm_oracle_run_calls.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_void_p),
                                           C.POINTER(_Stats), C.POINTER(C.c_uint64), C.c_uint32]
        L.hvm_oracle_run_calls.restype = C.c_int
        L.hvm_oracle_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def run_oracle(words, calls: bool = False) -> OracleResult:
    L = lib()
    arr = (C.c_uint32 * len(words))(*words)
    res = C.c_void_p()
    st = _Stats()
    ndefs = words[0] if words else 0
    per_fid = (C.c_uint64 * max(ndefs, 1))()
    if calls:
        rc = L.hvm_oracle_run_calls(arr, len(words), C.byref(res), C.byref(st), per_fid, ndefs)
    else:
        rc = L.hvm_oracle_run(arr, len(words), C.byref(res), C.byref(st))
    out = OracleResult(status=rc)
    if calls:
        out.calls = list(per_fid[:ndefs])
    if res.value:
        out.result = C.string_at(res.value).decode()
        L.hvm_oracle_free(res.value)
    out.itrs = st.itrs
    out.kinds = {k: st.kind[i] for i, k in enumerate(KINDS)}
    out.bytes = st.bytes
    out.normalizes = st.normalizes
    out.peak_nodes = st.peak_nodes
    out.seconds = st.seconds
    return out


@dataclass
class RefResult:
    returncode: int
    result: Optional[str]
    itrs: Optional[int]
    time_s: Optional[float]
    mips: Optional[float]
    stdout: str
    stderr: str


def ref_binary(threads: int) -> str:
    return os.path.join(REF_DIR, f"hvm_ref_t{threads}")


def have_reference(threads: int = 1) -> bool:
    return os.access(ref_binary(threads), os.X_OK)


def run_reference(words, threads: int = 8, timeout: float = 600.0) -> RefResult:
    """Runs the compiled reference run-c on a book buffer (child process)."""
    exe = ref_binary(threads)
    if not os.access(exe, os.X_OK):
        raise RuntimeError(f"{exe} missing: run `make -C oracle` where /root/reference exists")
    with tempfile.NamedTemporaryFile(suffix=".book", delete=False) as f:
        f.write((C.c_uint32 * len(words))(*words))
        path = f.name
    try:
        p = subprocess.run([exe, path], capture_output=True, text=True, timeout=timeout)
    finally:
        os.unlink(path)
    res = itrs = t = mips = None
    for line in p.stdout.splitlines():
        if line.startswith("Result: "):
            res = line[len("Result: "):]
        m = re.match(r"- ITRS: (\d+)", line)
        if m:
            itrs = int(m.group(1))
        m = re.match(r"- TIME: ([0-9.]+)s", line)
        if m:
            t = float(m.group(1))
        m = re.match(r"- MIPS: ([0-9.]+)", line)
        if m:
            mips = float(m.group(1))
    return RefResult(p.returncode, res, itrs, t, mips, p.stdout, p.stderr)


# ---------------------------------------------------------------------------
# Protocol model (hvm_model.cpp): the HIP turn kernel's steps executed on the
# CPU with the memory operations of all waves interleaved in a seeded random
# order.  It counts the events the protocol rules out (held vars read as free,
# takes that find no node, links that exceed the round cap, put-backs over a
# new node) and returns ITRS and the normal form for comparison with the
# oracle.  TEST INFRASTRUCTURE ONLY, like everything in this directory.
# ---------------------------------------------------------------------------

MODEL_LIB = os.path.join(HERE, "libhvm_model.so")
MODEL_XCHG_TWO = 1
MODEL_LATE_TAKES = 2
MODEL_OVERLAP_SCAN = 4
MODEL_TAKE_BEFORE_FIT = 8
MODEL_XCHG_ONE = 16
MODEL_NO_RECYCLE = 32
MODEL_RECYCLE_KEEP = 64
MODEL_VRECYCLE_KEEP = 128
MODEL_VRECYCLE = 256
MODEL_ENTER_LOAD = 512
MODEL_POOL = 1024


class _ModelCfg(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("lanes", C.c_uint32), ("deq", C.c_uint32), ("tbl", C.c_uint32),
                ("node_cap", C.c_uint32), ("vars_cap", C.c_uint32), ("root_vars", C.c_uint32),
                ("bag_cap", C.c_uint32), ("turn_steps", C.c_uint32), ("idle_pct", C.c_uint32),
                ("seed", C.c_uint32), ("switch_pct", C.c_uint32), ("variant", C.c_uint32),
                ("link_cap", C.c_uint32), ("max_ops", C.c_uint64)]


class _ModelOut(C.Structure):
    _fields_ = [("itrs", C.c_uint64), ("kind", C.c_uint64 * 10), ("err", C.c_uint32), ("turns", C.c_uint32),
                ("zero_publish", C.c_uint64), ("zero_enter", C.c_uint64), ("take_zero", C.c_uint64),
                ("putback", C.c_uint64), ("clobber", C.c_uint64), ("double_alloc", C.c_uint64),
                ("max_rounds", C.c_uint64), ("full_scans", C.c_uint64), ("pool_claims", C.c_uint64), ("defers", C.c_uint64),
                ("ops", C.c_uint64), ("first_event_op", C.c_uint64),
                ("first_event", C.c_uint32), ("first_err_turn", C.c_uint32), ("first_err_op", C.c_uint64),
                ("pool_gives", C.c_uint64), ("pool_takes", C.c_uint64), ("result", C.c_void_p)]


@dataclass
class ModelResult:
    status: int
    itrs: int
    kinds: dict
    err: int
    turns: int
    events: dict
    results: list


_mlib = None


def model_lib():
    global _mlib
    if _mlib is None:
        if not os.path.exists(MODEL_LIB):
            raise RuntimeError(f"{MODEL_LIB} missing: run `make -C oracle`")
        L = C.CDLL(MODEL_LIB)
        L.hvm_model_run.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_uint32), C.c_uint32,
                                    C.POINTER(_ModelCfg), C.POINTER(_ModelOut)]
        L.hvm_model_run.restype = C.c_int
        L.hvm_model_free.argtypes = [C.c_void_p]
        _mlib = L
    return _mlib


def run_model(words, entries=(0,), **cfg) -> ModelResult:
    """Runs the protocol model on a book; `entries` are the fids of the nets
    booted together (one: @main against ROOT; several: one root var each)."""
    L = model_lib()
    c = _ModelCfg()
    defaults = dict(waves=4, lanes=8, deq=8, tbl=256, node_cap=1 << 14, vars_cap=1 << 14, root_vars=0,
                    bag_cap=0, turn_steps=0, idle_pct=0, seed=0, switch_pct=100, variant=0, link_cap=0,
                    max_ops=0)
    defaults.update(cfg)
    for k, v in defaults.items():
        setattr(c, k, v)
    arr = (C.c_uint32 * len(words))(*words)
    ent = (C.c_uint32 * len(entries))(*entries)
    o = _ModelOut()
    rc = L.hvm_model_run(arr, len(words), ent, len(entries), C.byref(c), C.byref(o))
    res = []
    if o.result:
        res = C.string_at(o.result).decode().split(";")
        L.hvm_model_free(o.result)
    ev = {k: getattr(o, k) for k in ("zero_publish", "zero_enter", "take_zero", "putback", "clobber",
                                     "double_alloc", "max_rounds", "full_scans", "pool_claims", "defers", "ops", "first_event_op", "first_event",
                                     "first_err_op", "pool_gives", "pool_takes")}
    return ModelResult(rc, o.itrs, {k: o.kind[i] for i, k in enumerate(KINDS)}, o.err, o.turns, ev, res)
