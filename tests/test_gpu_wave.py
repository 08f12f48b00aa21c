"""The evaluator's cross-lane primitives (hip/hvm_wave.hpp: DPP row shifts
plus readlane) against numpy on the GPU: exclusive prefix sum and wave sum
(mod 2^32), wave max, and 64-bit wave sum (mod 2^64, carries between the
halves), on random inputs and on edge cases — zeros, all ones, a single
non-zero lane at every position, values that wrap.  The prefix sum is also
compared with the same operation over LDS permutes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M32, M64 = (1 << 32) - 1, (1 << 64) - 1


def cases():
    rng = np.random.default_rng(11)
    waves = []
    waves.append(([0] * 64, [0] * 64))
    waves.append(([M32] * 64, [M64] * 64))
    for pos in range(64):
        a = [0] * 64
        b = [0] * 64
        a[pos] = 0x80000001
        b[pos] = (1 << 63) | pos
        waves.append((a, b))
    for _ in range(40):
        waves.append((rng.integers(0, 1 << 32, 64, dtype=np.uint64).tolist(),
                      rng.integers(0, 1 << 63, 64, dtype=np.uint64).tolist()))
    for _ in range(20):   # small counts, as the kernels scan them
        waves.append((rng.integers(0, 300, 64).tolist(), rng.integers(0, 1 << 40, 64).tolist()))
    return waves


def test_wave_primitives_match_numpy(hip):
    waves = cases()
    in32 = [int(x) for a, _ in waves for x in a]
    in64 = [int(x) for _, b in waves for x in b]
    out = hip.wave_selftest(in32, in64)
    for w, (a, b) in enumerate(waves):
        excl, acc = [], 0
        for x in a:
            excl.append(acc & M32)
            acc += int(x)
        total, mx, s64 = acc & M32, max(int(x) for x in a), sum(int(x) for x in b) & M64
        for lane in range(64):
            sc, tot, m, sc_ref, tot_ref, s = out[w * 64 + lane]
            assert (sc, tot, m, s) == (excl[lane], total, mx, s64), (w, lane)
            assert (sc_ref, tot_ref) == (sc, tot), (w, lane)
