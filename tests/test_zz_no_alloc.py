"""hvm_hip_run_static: the entry point that allocates nothing (its net, bags,
control blocks, book and readback lists live in the library's static device
array), called through the C ABI and checked against hvm_hip_run, the entry
point it is a variant of: same ITRS, per-kind counts and normal form, on two
calls in a row, the second with a smaller input (the array keeps the first
call's contents, which must not leak into the second).  Requests beyond the
array's limits are refused with HVM_HIP_E_ARG and a reason."""
import ctypes as C

import pytest

import programs as P

pytestmark = pytest.mark.gpu

E_ARG = 5


def run_static(hip, words, result_cap=1 << 20, **cfg):
    L = hip.lib()
    f = L.hvm_hip_run_static
    f.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(hip._Cfg), C.POINTER(hip._Out),
                  C.c_char_p, C.c_size_t]
    f.restype = C.c_int
    buf = C.create_string_buffer(result_cap) if result_cap else None
    out = hip._Out()
    c = hip._cfg(**cfg)
    arr = hip.words_array(words)
    rc = f(arr, len(words), C.byref(c), C.byref(out), buf, result_cap)
    text = buf.value.decode() if buf is not None else None
    return rc, out, text


def test_static_entry_matches_hvm_hip_run(hip):
    cases = [P.radix_lsd(10, 3)[0], P.bitonic(6, 2)[0]]   # the second call is the smaller one
    for code in cases:
        words = hip.compile_book(code)
        ref = hip.run_hip(words, time_limit_s=120)
        print("static entry: itrs", ref.itrs, flush=True)
        rc, out, text = run_static(hip, words, time_limit_s=120)
        assert rc == 0, hip.lib().hvm_hip_last_error(None)
        assert out.err == 0
        assert out.itrs == ref.itrs
        assert [out.kind[i] for i in range(10)] == [ref.kinds[k] for k in hip.KINDS]
        assert out.bytes == ref.bytes
        assert hip.canonical(text) == ref.canonical
        assert not out.result and not out.canonical       # the text went to the caller's buffer


def test_static_entry_refuses_what_does_not_fit(hip):
    words = hip.compile_book(P.sum_tree(4)[0])
    rc, _, _ = run_static(hip, words, node_cap=1 << 22)
    assert rc == E_ARG
    assert b"HVM_HIP_STATIC_SLOTS" in hip.lib().hvm_hip_last_error(None)
    rc, _, _ = run_static(hip, words, flags=hip.GRAPHS)
    assert rc == E_ARG
    rc, _, _ = run_static(hip, words, result_cap=2)    # "16" needs 3 bytes
    assert rc == E_ARG
    # and the array still works after the refusals
    rc, out, text = run_static(hip, words)
    assert rc == 0 and text == "16"
