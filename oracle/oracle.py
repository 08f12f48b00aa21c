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
