"""Python host for the MI355X HVM2 evaluator (ctypes over lib/libhvm_hip.so).

Mirrors the reference's host interface for the GPU path: `compile(code)` is
`ast::Book::parse(..).build().to_buffer(..)` (src/main.rs:87-92) and
`run_hip(words)` is the call into the device runtime (`hvm_cu`, src/main.rs:93-96),
returning the result text and counters instead of printing them.  The library
must be built (`make -C lambdavm-rust_amd`); there is no CPU fallback: if the
shared library or the GPU is missing, these functions raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
# HVM_HIP_LIB selects another build of the library, such as a variant with
# other compile-time knobs (lib/variants/, searched by autotune.py).
LIB_PATH = os.environ.get("HVM_HIP_LIB") or os.path.join(_HERE, "lib", "libhvm_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "hvm_hip.h")

KINDS = ["CALL", "CALL_COPY", "VOID", "ERAS", "ANNI", "COMM", "OPER1", "OPER2", "SWIT0", "SWITN"]

# Algorithmic bytes per interaction kind (CALL is per def: 8N + 4V + 8(R+1)).
KIND_BYTES = {"CALL_COPY": 32, "VOID": 0, "ERAS": 32, "ANNI": 48, "COMM": 112,
              "OPER1": 32, "OPER2": 40, "SWIT0": 32, "SWITN": 40}

NO_READBACK = 1
FENCE_RELEASE = 2
PROFILE = 4
LOCAL_NET = 8
COMPILED_CALLS = 16
WIDE = 32
GRAPHS = 64
SPEC_NODES = 64

# Scheduler (include/hvm_hip.h)
SCHED_PLAIN, SCHED_LEARNED, SCHED_EXPLORE, SCHED_FIXED = 0, 1, 2, 3
TURN_LENGTHS = (4096, 4096, 4096, 512)
IDLE_PCTS = (10, None, 50, None)   # None: the configured idle_pct
MLP_FEAT, MLP_HID, MLP_ACT = 32, 64, 4
MLP_PARAMS = MLP_FEAT * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_ACT + MLP_ACT
TRACE_REC = MLP_FEAT + 4

STATUS = {0: "ok", 1: "book failed to load", 2: "IO requested (unsupported)",
          3: "device error", 4: "evaluator error", 5: "bad argument"}


class HvmError(RuntimeError):
    pass


class HvmEvalError(HvmError):
    """The evaluator stopped: error flags (`err`, HVM_ERR_* bits) or its time
    limit.  Context creation and book loading raise plain HvmError."""

    def __init__(self, msg: str, err: int = 0):
        super().__init__(msg)
        self.err = err


class _Cfg(C.Structure):
    _fields_ = [("node_cap", C.c_uint64), ("vars_cap", C.c_uint64), ("bag_cap", C.c_uint32),
                ("waves", C.c_uint32), ("turn_steps", C.c_uint32), ("flags", C.c_uint32),
                ("time_limit_s", C.c_uint32), ("idle_pct", C.c_uint32)]


class _Out(C.Structure):
    _fields_ = [("itrs", C.c_uint64), ("kind", C.c_uint64 * 10), ("bytes", C.c_uint64),
                ("leak", C.c_uint64), ("steps", C.c_uint64), ("spills", C.c_uint64),
                ("defers", C.c_uint64), ("retries", C.c_uint64), ("probes", C.c_uint64),
                ("kernel_ms", C.c_double), ("eval_ms", C.c_double), ("wall_s", C.c_double),
                ("turns", C.c_uint32), ("normalizes", C.c_uint32), ("err", C.c_uint32),
                ("launches", C.c_uint32), ("launch_ms", C.c_double), ("waves", C.c_uint32),
                ("saves", C.c_uint64), ("local_left", C.c_uint64),
                ("phase_ms", C.c_double * 6),
                ("result", C.c_void_p), ("canonical", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    """Loads lib/libhvm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HvmError(f"{LIB_PATH} is missing: build it with `make -C lambdavm-rust_amd`")
    L = C.CDLL(LIB_PATH)
    L.hvm_front_compile.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_uint32)),
                                    C.POINTER(C.c_size_t), C.POINTER(C.c_void_p)]
    L.hvm_front_compile.restype = C.c_int
    L.hvm_front_canonical.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    L.hvm_front_canonical.restype = C.c_int
    L.hvm_front_specialize.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.c_uint32,
                                       C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_size_t),
                                       C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_void_p)]
    L.hvm_front_specialize.restype = C.c_int
    L.hvm_hip_free.argtypes = [C.c_void_p]
    L.hvm_hip_free.restype = None
    L.hvm_hip_create.argtypes = [C.POINTER(_Cfg)]
    L.hvm_hip_create.restype = C.c_void_p
    L.hvm_hip_load.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.c_size_t]
    L.hvm_hip_load.restype = C.c_int
    L.hvm_hip_eval.argtypes = [C.c_void_p, C.POINTER(_Out)]
    L.hvm_hip_eval.restype = C.c_int
    L.hvm_hip_destroy.argtypes = [C.c_void_p]
    L.hvm_hip_destroy.restype = None
    L.hvm_hip_last_error.argtypes = [C.c_void_p]
    L.hvm_hip_last_error.restype = C.c_char_p
    L.hvm_hip_run.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(_Cfg), C.POINTER(_Out)]
    L.hvm_hip_run.restype = C.c_int
    L.hvm_hip_eval_batch.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32,
                                     C.POINTER(C.c_uint64), C.POINTER(_Out), C.POINTER(_Out)]
    L.hvm_hip_eval_batch.restype = C.c_int
    L.hvm_hip_set_scheduler.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_uint16), C.c_uint32]
    L.hvm_hip_set_scheduler.restype = C.c_int
    L.hvm_hip_read_trace.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_uint32)]
    L.hvm_hip_read_trace.restype = C.c_int
    F = C.POINTER(C.c_float)
    U = C.POINTER(C.c_uint32)
    H = C.POINTER(C.c_uint16)
    L.hvm_mlp_forward.argtypes = [H, F, C.c_uint32, F]
    L.hvm_mlp_forward.restype = C.c_int
    L.hvm_mlp_grad.argtypes = [H, F, U, F, C.c_uint32, F, F]
    L.hvm_mlp_grad.restype = C.c_int
    L.hvm_mlp_train.argtypes = [F, F, U, F, C.c_uint32, C.c_uint32, C.c_float, H, F, F]
    L.hvm_mlp_train.restype = C.c_int
    _lib = L
    return L


def _take_str(p) -> Optional[str]:
    if not p:
        return None
    s = C.string_at(p).decode("utf-8", "replace")
    lib().hvm_hip_free(p)
    return s


def compile_book(code: str) -> List[int]:
    """`.hvm` text → Book::to_buffer words (list of u32)."""
    L = lib()
    words = C.POINTER(C.c_uint32)()
    n = C.c_size_t()
    err = C.c_void_p()
    rc = L.hvm_front_compile(code.encode(), C.byref(words), C.byref(n), C.byref(err))
    if rc != 0:
        raise HvmError(_take_str(err.value) or "compile failed")
    out = words[: n.value]
    L.hvm_hip_free(C.cast(words, C.c_void_p))
    return list(out)


def specialize(words, node_budget: int = SPEC_NODES):
    """The load-time specialisation of COMPILED_CALLS on a wire buffer:
    returns (specialised words, [(static CALLs, static ANNIs, bytes)] per def)."""
    L = lib()
    arr = words_array(words)
    out = C.POINTER(C.c_uint32)()
    n = C.c_size_t()
    cnt = C.POINTER(C.c_uint32)()
    err = C.c_void_p()
    rc = L.hvm_front_specialize(arr, len(words), node_budget, C.byref(out), C.byref(n), C.byref(cnt), C.byref(err))
    if rc != 0:
        raise HvmError(_take_str(err.value) or f"specialize failed ({rc})")
    new = list(out[: n.value])
    ndefs = new[0]
    counts = [tuple(cnt[3 * f: 3 * f + 3]) for f in range(ndefs)]
    L.hvm_hip_free(C.cast(out, C.c_void_p))
    L.hvm_hip_free(C.cast(cnt, C.c_void_p))
    return new, counts


def canonical(printed: str) -> str:
    """Canonical form of a printed tree (vars renamed x0, x1, ... in order)."""
    L = lib()
    out = C.c_void_p()
    err = C.c_void_p()
    if L.hvm_front_canonical(printed.encode(), C.byref(out), C.byref(err)) != 0:
        raise HvmError(_take_str(err.value) or "canonical failed")
    return _take_str(out.value)


def words_array(words) -> C.Array:
    return (C.c_uint32 * len(words))(*words)


@dataclass
class RunResult:
    status: int
    itrs: int = 0
    kinds: dict = field(default_factory=dict)
    bytes: int = 0
    leak: int = 0
    kernel_ms: float = 0.0
    eval_ms: float = 0.0
    wall_s: float = 0.0
    turns: int = 0
    normalizes: int = 0
    err: int = 0
    launches: int = 0
    launch_ms: float = 0.0
    result: Optional[str] = None
    canonical: Optional[str] = None
    diag: dict = field(default_factory=dict)


def _cfg(node_cap=0, vars_cap=0, bag_cap=0, waves=0, turn_steps=0, flags=0, time_limit_s=0,
         idle_pct=0) -> _Cfg:
    return _Cfg(node_cap, vars_cap, bag_cap, waves, turn_steps, flags, time_limit_s, idle_pct)


def _result(rc: int, o: _Out) -> RunResult:
    r = RunResult(status=rc, itrs=o.itrs, kinds={k: o.kind[i] for i, k in enumerate(KINDS)},
                  bytes=o.bytes, leak=o.leak, kernel_ms=o.kernel_ms, eval_ms=o.eval_ms,
                  wall_s=o.wall_s, turns=o.turns, normalizes=o.normalizes, err=o.err,
                  launches=o.launches, launch_ms=o.launch_ms,
                  diag={"steps": o.steps, "spills": o.spills, "defers": o.defers,
                        "retries": o.retries, "probes": o.probes, "waves": o.waves,
                        "saves": o.saves, "local_left": o.local_left,
                        "phase_ms": [round(o.phase_ms[k], 3) for k in range(6)]})
    r.result = _take_str(o.result)
    r.canonical = _take_str(o.canonical)
    return r


class Context:
    """A persistent evaluator: net, bags and book stay in HBM between runs."""

    def __init__(self, **cfg):
        L = lib()
        self._cfg = _cfg(**cfg)
        self._h = L.hvm_hip_create(C.byref(self._cfg))
        if not self._h:
            raise HvmError("hvm_hip_create failed: " + L.hvm_hip_last_error(None).decode())
        self._words = None

    def load(self, words) -> None:
        self._words = words_array(words)
        rc = lib().hvm_hip_load(self._h, self._words, len(words))
        if rc != 0:
            raise HvmError(f"load failed ({STATUS.get(rc, rc)}): " + self.last_error())

    def eval(self) -> RunResult:
        o = _Out()
        rc = lib().hvm_hip_eval(self._h, C.byref(o))
        r = _result(rc, o)
        if rc != 0:
            msg = f"eval failed ({STATUS.get(rc, rc)}, err=0x{r.err:x}): " + self.last_error()
            raise (HvmEvalError(msg, r.err) if rc == 4 else HvmError(msg))
        return r

    def eval_batch(self, entries, est_nodes=None):
        """Evaluates independent nets of the loaded book (hvm_hip_eval_batch).
        Returns (per-net results, total)."""
        L = lib()
        n = len(entries)
        ents = (C.c_uint32 * n)(*entries)
        est = (C.c_uint64 * n)(*est_nodes) if est_nodes is not None else None
        outs = (_Out * n)()
        total = _Out()
        rc = L.hvm_hip_eval_batch(self._h, ents, n, est, outs, C.byref(total))
        per = [_result(outs[i].err, outs[i]) for i in range(n)]
        tot = _result(rc, total)
        if rc != 0:
            msg = f"eval_batch failed ({STATUS.get(rc, rc)}, err=0x{tot.err:x}): " + self.last_error()
            raise (HvmEvalError(msg, tot.err) if rc == 4 else HvmError(msg))
        return per, tot

    def set_scheduler(self, mode: int = SCHED_PLAIN, seed: int = 0, explore_pct: int = 0,
                      params_fp16=None, trace_cap: int = 0) -> None:
        """Selects the turn-length scheduler (see include/hvm_hip.h)."""
        p = None
        if params_fp16 is not None:
            p = (C.c_uint16 * MLP_PARAMS)(*[int(v) for v in params_fp16])
        rc = lib().hvm_hip_set_scheduler(self._h, mode, seed, explore_pct, p, trace_cap)
        if rc != 0:
            raise HvmError("set_scheduler failed: " + self.last_error())

    def read_trace(self):
        """Per-decision records: list of (features[32], action, reward, itrs, us)."""
        L = lib()
        n = C.c_uint32()
        L.hvm_hip_read_trace(self._h, None, 0, C.byref(n))
        if n.value == 0:
            return []
        buf = (C.c_float * (n.value * TRACE_REC))()
        rc = L.hvm_hip_read_trace(self._h, buf, n.value, C.byref(n))
        if rc != 0:
            raise HvmError("read_trace failed: " + self.last_error())
        out = []
        for i in range(n.value):
            r = buf[i * TRACE_REC:(i + 1) * TRACE_REC]
            out.append((list(r[:MLP_FEAT]), int(r[MLP_FEAT]), r[MLP_FEAT + 1], r[MLP_FEAT + 2], r[MLP_FEAT + 3]))
        return out

    def last_error(self) -> str:
        return lib().hvm_hip_last_error(self._h).decode()

    def close(self) -> None:
        if self._h:
            lib().hvm_hip_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_hip(words, **cfg) -> RunResult:
    """One-shot evaluation of a book buffer on the GPU (hvm_hip_run)."""
    L = lib()
    arr = words_array(words)
    c = _cfg(**cfg)
    o = _Out()
    rc = L.hvm_hip_run(arr, len(words), C.byref(c), C.byref(o))
    r = _result(rc, o)
    if rc != 0:
        msg = f"hvm_hip_run failed: {STATUS.get(rc, rc)} (err=0x{r.err:x})"
        raise (HvmEvalError(msg, r.err) if rc == 4 else HvmError(msg))
    return r


def run_code(code: str, **cfg) -> RunResult:
    return run_hip(compile_book(code), **cfg)


# ---------------------------------------------------------------------------
# Scheduler MLP (training runs on the GPU through the library's HIP kernels)
# ---------------------------------------------------------------------------

def _farr(xs):
    return (C.c_float * len(xs))(*xs)


def atomic_rate(slots: int, nbytes: int) -> float:
    """Device-scope exchanges per second on random slots (hvm_hip_atomic_rate)."""
    L = lib()
    L.hvm_hip_atomic_rate.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
    out = C.c_double(0)
    rc = L.hvm_hip_atomic_rate(slots, nbytes, C.byref(out))
    if rc != 0:
        raise HvmError(f"hvm_hip_atomic_rate failed: {STATUS.get(rc, rc)}")
    return out.value


def copy_rate(nbytes: int) -> float:
    """Bytes read plus written per second by a device-to-device copy of
    `nbytes` (hvm_hip_copy_rate, best of 5)."""
    L = lib()
    L.hvm_hip_copy_rate.argtypes = [C.c_uint64, C.POINTER(C.c_double)]
    out = C.c_double(0)
    rc = L.hvm_hip_copy_rate(nbytes, C.byref(out))
    if rc != 0:
        raise HvmError(f"hvm_hip_copy_rate failed: {STATUS.get(rc, rc)}")
    return out.value


def wave_selftest(in32, in64):
    """hvm_hip_wave_selftest on len(in32) // 64 waves: per lane (scan, total,
    max, scan over LDS permutes, their total, 64-bit sum)."""
    n = len(in32)
    assert n % 64 == 0 and len(in64) == n
    L = lib()
    L.hvm_hip_wave_selftest.argtypes = [C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint32)]
    out = (C.c_uint32 * (n * 7))()
    rc = L.hvm_hip_wave_selftest(n // 64, (C.c_uint32 * n)(*in32), (C.c_uint64 * n)(*in64), out)
    if rc != 0:
        raise HvmError(f"hvm_hip_wave_selftest failed: {STATUS.get(rc, rc)}")
    o = list(out)
    return [(o[7 * i], o[7 * i + 1], o[7 * i + 2], o[7 * i + 3], o[7 * i + 4], o[7 * i + 5] | (o[7 * i + 6] << 32))
            for i in range(n)]


def request_rate(slots: int, nbytes: int, op: int) -> float:
    """Device-scope requests per second on random slots (hvm_hip_request_rate):
    op 0 exchanges, 1 stores, 2 loads."""
    L = lib()
    L.hvm_hip_request_rate.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]
    out = C.c_double(0)
    rc = L.hvm_hip_request_rate(slots, nbytes, op, C.byref(out))
    if rc != 0:
        raise HvmError(f"hvm_hip_request_rate failed: {STATUS.get(rc, rc)}")
    return out.value


def mlp_forward(params_fp16, x_rows):
    """x_rows: list of 32-float rows → list of 4-float predictions."""
    n = len(x_rows)
    p = (C.c_uint16 * MLP_PARAMS)(*params_fp16)
    x = _farr([v for row in x_rows for v in row])
    y = (C.c_float * (n * MLP_ACT))()
    rc = lib().hvm_mlp_forward(p, x, n, y)
    if rc != 0:
        raise HvmError(f"hvm_mlp_forward failed: {STATUS.get(rc, rc)}")
    return [list(y[i * MLP_ACT:(i + 1) * MLP_ACT]) for i in range(n)]


def mlp_grad(params_fp16, x_rows, actions, rewards):
    """Mean gradient (packed like the params) and mean loss, on the GPU."""
    n = len(x_rows)
    p = (C.c_uint16 * MLP_PARAMS)(*params_fp16)
    x = _farr([v for row in x_rows for v in row])
    a = (C.c_uint32 * n)(*actions)
    r = _farr(rewards)
    g = (C.c_float * MLP_PARAMS)()
    loss = C.c_float()
    rc = lib().hvm_mlp_grad(p, x, a, r, n, g, C.byref(loss))
    if rc != 0:
        raise HvmError(f"hvm_mlp_grad failed: {STATUS.get(rc, rc)}")
    return list(g), loss.value


def mlp_train(params_fp32, x_rows, actions, rewards, epochs=300, lr=3e-3):
    """Full-batch Adam on the GPU.  Returns (fp16 params, fp32 params, loss)."""
    n = len(x_rows)
    pin = _farr(params_fp32)
    x = _farr([v for row in x_rows for v in row])
    a = (C.c_uint32 * n)(*actions)
    r = _farr(rewards)
    out16 = (C.c_uint16 * MLP_PARAMS)()
    out32 = (C.c_float * MLP_PARAMS)()
    loss = C.c_float()
    rc = lib().hvm_mlp_train(pin, x, a, r, n, epochs, lr, out16, out32, C.byref(loss))
    if rc != 0:
        raise HvmError(f"hvm_mlp_train failed: {STATUS.get(rc, rc)}")
    return list(out16), list(out32), loss.value
