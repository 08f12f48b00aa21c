"""The HIP turn kernel's protocol, run on the CPU by the protocol model
(oracle/hvm_model.cpp): the kernel's steps with the memory operations of all
waves interleaved in a seeded random order.

The batched workload with idle quit off is the case that raised the kernel's
link-loop flag (0x8) on the GPU: few busy waves run long turns, fill their
node and var slices and allocate from the shared pool.  Its cause, found with
this model: an allocation that scanned its whole slice went on for one more
window, wrapped around onto the slots it had handed out earlier in the same
call (still zero until the creation phase) and handed them out again.  Two
interactions then shared a var slot, a held var read as free, and a link
published into slot 0 (ROOT's slot, zero in a batched evaluation) forever.

Checks: with the kernel's current protocol no such event occurs on any seed
and every net's normal form and the total ITRS equal the oracle's; with the
old scan planted back into the model, the events and the flag come back (so
the model does see them).  The batch is a small replica of the 256-net
workload (sum trees and bitonic sorts, one root var each).
"""
import pytest

import programs as P

ANOMALIES = ("zero_publish", "zero_enter", "take_zero", "putback", "clobber", "double_alloc")


def mini_batch(seed, n, dmin=4, dmax=9, kmin=2, kmax=5):
    g = P.splitmix64(seed ^ 0x5EED_BA7C)
    body = ["@main = *", P.SUM_DEFS.strip(), P.BITONIC_DEFS.strip()]
    nets = []
    for i in range(n):
        if i < n // 2:
            d = dmin + next(g) % (dmax - dmin + 1)
            body.append(f"@net{i} = r & @sum ~ ({d} r)")
            nets.append((P.sum_tree(d)[0], str((1 << d) % (1 << 24))))
        else:
            k = kmin + next(g) % (kmax - kmin + 1)
            ks = next(g) & 0xFFFFFFFF
            keys = P.keys_uniform(1 << k, ks)
            top, lines = P.key_tree_defs(keys, f"k{i}", 3)
            body.append(f"@net{i} = r & @sort ~ ({k} (0 (@{top} r)))")
            body.extend(lines)
            nets.append((P.bitonic(k, ks, 3)[0], sorted(keys)))
    code = "\n".join(body) + "\n"
    fids = P.def_fids(code)
    return code, [fids[f"net{i}"] for i in range(n)], nets


def result_ok(text, expect):
    if isinstance(expect, list):
        try:
            return P.flatten_tree(P.parse_tree_text(text)) == expect
        except Exception:
            return False
    return text == expect


# 16 waves of 8 lanes, 16K-slot net: slices fill, the pool is used
CFG = dict(waves=16, lanes=8, deq=8, node_cap=1 << 14, vars_cap=1 << 14, root_vars=4097, idle_pct=101,
           max_ops=20_000_000)


@pytest.fixture(scope="module")
def batch(hip, orc):
    code, entries, nets = mini_batch(0, 32)
    words = hip.compile_book(code)
    itrs = sum(orc.run_oracle(hip.compile_book(c)).itrs for c, _ in nets)
    return words, entries, nets, itrs


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_batch_idle_off_is_clean(batch, orc, seed):
    words, entries, nets, itrs = batch
    r = orc.run_model(words, entries, seed=seed, **CFG)
    assert r.status == 0 and r.err == 0
    assert {k: r.events[k] for k in ANOMALIES} == {k: 0 for k in ANOMALIES}
    assert r.events["full_scans"] > 0 and r.events["pool_claims"] > 0   # the case is exercised
    assert r.itrs == itrs
    assert all(result_ok(t, e) for t, (_, e) in zip(r.results, nets))


@pytest.mark.parametrize("variant", ["MODEL_XCHG_TWO", "MODEL_XCHG_ONE", "MODEL_NO_RECYCLE", "MODEL_ENTER_LOAD"])
def test_batch_idle_off_exchange_takes_is_clean(batch, orc, variant):
    # the earlier take modes (exchanges, put back on a retry): of ANNI/COMM,
    # and of single-take rules and a CALL's copy (HVM_PEEK_TAKES=0); node
    # slots zeroed at commit (HVM_NODE_RECYCLE=0); ENTER by load
    # (HVM_ENTER_LOAD=1)
    words, entries, nets, itrs = batch
    r = orc.run_model(words, entries, seed=3, variant=getattr(orc, variant), **CFG)
    assert r.err == 0 and r.itrs == itrs
    assert r.events["take_zero"] == 0


@pytest.mark.parametrize("variant", ["MODEL_RECYCLE_KEEP", "MODEL_VRECYCLE_KEEP"])
def test_recycled_slot_handed_out_twice_is_detected(batch, orc, variant):
    # a planted mistake: slots taken from the wave's node (var) recycling
    # stack stay on it, so a later allocation hands them out again while they
    # are in use
    words, entries, nets, itrs = batch
    r = orc.run_model(words, entries, seed=0, variant=getattr(orc, variant) | orc.MODEL_VRECYCLE, **CFG)
    assert r.events["double_alloc"] > 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mid_turn_pool_is_clean(batch, orc, seed):
    # the mid-turn pool (a build with HVM_MID_POOL): busy waves hand surplus
    # redexes to the output bag, waves that run dry take them back by
    # exchanges with 0, and the stores of handed-over redexes land several
    # scheduling points after their reservation.  On a net 4 times larger:
    # on the 16K-slot one the pool's extra concurrency runs out of var slots
    # at most seeds (a capacity effect: the larger net passes every seed).
    words, entries, nets, itrs = batch
    cfg = dict(CFG, node_cap=1 << 16, vars_cap=1 << 16)
    r = orc.run_model(words, entries, seed=seed, variant=orc.MODEL_POOL, **cfg)
    assert r.status == 0 and r.err == 0
    assert {k: r.events[k] for k in ANOMALIES} == {k: 0 for k in ANOMALIES}
    assert r.events["pool_gives"] > 0 and r.events["pool_takes"] > 0   # the pool is exercised
    assert r.itrs == itrs
    assert all(result_ok(t, e) for t, (_, e) in zip(r.results, nets))


@pytest.mark.parametrize("seed", [4, 5])
def test_batch_with_var_recycling_is_clean(batch, orc, seed):
    # var slots freed by links recycled as well (a build with HVM_VAR_RECYCLE)
    words, entries, nets, itrs = batch
    r = orc.run_model(words, entries, seed=seed, variant=orc.MODEL_VRECYCLE, **CFG)
    assert r.status == 0 and r.err == 0
    assert {k: r.events[k] for k in ANOMALIES} == {k: 0 for k in ANOMALIES}
    assert r.itrs == itrs
    assert all(result_ok(t, e) for t, (_, e) in zip(r.results, nets))


def test_overlapping_scan_is_detected(batch, orc):
    # the old allocator scan (len + one window) planted back into the model
    words, entries, nets, itrs = batch
    r = orc.run_model(words, entries, seed=0, variant=orc.MODEL_OVERLAP_SCAN, **CFG)
    assert r.err & 0x8                       # HVM_ERR_LINK_LOOP, as on the GPU
    assert r.events["zero_publish"] + r.events["zero_enter"] > 0
    assert r.events["first_event_op"] < r.events["first_err_op"]   # the anomaly comes first


@pytest.mark.parametrize("gen,args", [("sum_tree", (8,)), ("bitonic", (5, 1)), ("radix", (4, 2)),
                                      ("church", (3, 4))])
def test_single_nets_match_oracle(hip, orc, gen, args):
    code, _ = getattr(P, gen)(*args)
    w = hip.compile_book(code)
    o = orc.run_oracle(w)
    for seed in range(2):
        r = orc.run_model(w, seed=seed, waves=4, lanes=8, node_cap=1 << 15, vars_cap=1 << 15, turn_steps=5)
        assert r.err == 0 and r.itrs == o.itrs and r.kinds == o.kinds
        assert hip.canonical(r.results[0]) == hip.canonical(o.result)


def test_takes_only_after_the_step_fits(hip, orc):
    """A lane takes its nodes only once its step is known to fit the
    allocation tables.  Planted in the model: every lane takes first and a lane
    that then has to wait puts its nodes back (the local-net evaluator's
    early-take variant did this); meanwhile the owning wave's allocator can
    hand out a slot that reads zero, and the put-back overwrites the new node.
    The model sees those overwrites, and wrong results follow; the kernel's
    protocol has none on the same schedule.  A COMM-heavy program on 64-lane
    waves with 16-entry tables makes lanes wait often."""
    code, expect = P.church_loop(6, 8)
    w = hip.compile_book(code)
    cfg = dict(waves=4, lanes=64, deq=16, tbl=16, node_cap=1 << 13, vars_cap=1 << 13, max_ops=30_000_000)
    ok = orc.run_model(w, seed=0, **cfg)
    assert ok.err == 0 and ok.results == [str(expect)] and ok.events["defers"] > 0
    assert {k: ok.events[k] for k in ANOMALIES} == {k: 0 for k in ANOMALIES}
    bad = orc.run_model(w, seed=0, variant=orc.MODEL_TAKE_BEFORE_FIT, **cfg)
    assert bad.events["clobber"] > 0
    assert bad.results != [str(expect)]
