"""GPU tests of the scheduler MLP kernels, the scheduler modes and batching.

* MLP forward and gradient (HIP, fp16 weights, fp32 math) against the float32
  numpy reference in tests/mlp_ref.py (itself checked against finite
  differences in test_sched_cpu.py); tolerance rtol 2e-3 / atol 2e-5;
* every scheduler mode (plain, explore with several seeds, fixed actions,
  learned with random and with trained weights) gives the same normal form
  and ITRS; the trained weights beat the plain heuristic at held-out seeds;
* the batched workload: every net's normal form equals the reference run-c
  result of that net alone (tests/golden/batch.json), and the total ITRS is
  their sum.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import mlp_ref
import programs as P
import sched as S
from conftest import ROOT

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2e-3, 2e-5


def rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.25, sum(mlp_ref.SIZES)).astype(np.float32)
    bits = mlp_ref.to_half_bits(p)
    X = rng.random((n, mlp_ref.F)).astype(np.float32)
    A = rng.integers(0, mlp_ref.K, n).astype(np.uint32)
    R = rng.normal(0, 1, n).astype(np.float32)
    return bits, X, A, R


def test_mlp_forward_matches_reference(hip):
    bits, X, A, R = rand_problem(300, 1)
    y = np.array(hip.mlp_forward(bits.tolist(), X.tolist()), dtype=np.float32)
    ref, _ = mlp_ref.forward(mlp_ref.from_half_bits(bits), X)
    assert np.allclose(y, ref, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("n,seed", [(1, 2), (7, 3), (513, 4)])
def test_mlp_gradient_matches_reference(hip, n, seed):
    bits, X, A, R = rand_problem(n, seed)
    g, loss = hip.mlp_grad(bits.tolist(), X.tolist(), A.tolist(), R.tolist())
    ref_loss, ref_g = mlp_ref.loss_and_grad(mlp_ref.from_half_bits(bits), X, A, R)
    assert np.allclose(np.array(g, dtype=np.float32), ref_g, rtol=RTOL, atol=ATOL)
    assert abs(loss - ref_loss) <= RTOL * max(1.0, abs(ref_loss))


def test_mlp_training_lowers_loss(hip):
    rng = np.random.default_rng(5)
    X = rng.random((256, mlp_ref.F)).astype(np.float32)
    A = rng.integers(0, mlp_ref.K, 256)
    R = (X[:, 0] * 2 - X[:, 1] + 0.5 * A).astype(np.float32)   # learnable target
    p0 = S.init_params(0)
    _, l0 = hip.mlp_grad(mlp_ref.to_half_bits(p0).tolist(), X.tolist(), A.tolist(), R.tolist())
    p16, p32, l1 = hip.mlp_train(p0, X.tolist(), A.tolist(), R.tolist(), epochs=300, lr=3e-3)
    assert l1 < 0.25 * l0


def test_scheduler_never_changes_results(hip):
    code, expect = P.bitonic(9, 4)
    words = hip.compile_book(code)
    seen = set()
    with hip.Context(time_limit_s=120) as ctx:
        ctx.load(words)
        runs = [(hip.SCHED_PLAIN, 0, 0, None)] + [(hip.SCHED_EXPLORE, s, 50, None) for s in (1, 2, 3)]
        # every action held fixed
        runs += [(hip.SCHED_FIXED, 7, a, None) for a in range(4)]
        rnd = mlp_ref.to_half_bits(np.random.default_rng(9).normal(0, 0.3, sum(mlp_ref.SIZES))).tolist()
        runs.append((hip.SCHED_LEARNED, 0, 0, rnd))
        for mode, seed, pct, params in runs:
            ctx.set_scheduler(mode, seed=seed, explore_pct=pct, params_fp16=params, trace_cap=4096)
            r = ctx.eval()
            assert P.flatten_tree(P.parse_tree_text(r.result)) == expect
            seen.add((r.itrs, r.canonical))
            tr = ctx.read_trace()
            assert len(tr) >= 1
            assert all(0 <= rec[1] < 4 for rec in tr)
            if mode == hip.SCHED_FIXED:
                assert all(rec[1] == pct for rec in tr)
    assert len(seen) == 1


def _tuned_batch_cfg():
    path = os.path.join(ROOT, "lambdavm-rust_amd", "tuned_batch.json")
    return json.load(open(path))["config"] if os.path.exists(path) else None


@pytest.mark.parametrize("which", ["default", "tuned", "idle_off", "idle_off_512"])
def test_batch_matches_per_net_reference(hip, which):
    """idle_off: turns never end early (idle_pct above 100), so few busy waves
    run long turns, fill their slices and allocate from the shared pool; the
    case that raised the link-loop flag before the allocator's scan stopped
    revisiting its first window (tests/test_model.py holds the CPU model of
    it).  idle_off_512: the same on a quarter of the default wave count."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "batch.json")))
    code, nets = P.batch(g["seed"], g["n"])
    fids = P.def_fids(code)
    entries = [fids[f"net{i}"] for i in range(len(nets))]
    est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
    cfg = {"node_cap": 1 << 29, "vars_cap": 1 << 29}
    if which.startswith("idle_off"):
        cfg["idle_pct"] = 101
        if which.endswith("512"):
            cfg["waves"] = 512
    if which == "tuned":
        if _tuned_batch_cfg() is None:
            pytest.skip("no tuned configuration committed")
        cfg = dict(_tuned_batch_cfg())
        cfg["flags"] = cfg.get("flags", 0) | cfg.pop("graphs", 0)   # the autotuner's graphs knob
        if cfg.pop("variant", None):
            pytest.skip("the tuned configuration uses a variant build: test_tuned_configurations_keep_parity")
    with hip.Context(time_limit_s=300, **cfg) as ctx:
        ctx.load(hip.compile_book(code))
        per, total = ctx.eval_batch(entries, est)
    assert total.itrs == g["total_itrs"]
    for i, (row, r) in enumerate(zip(g["nets"], per)):
        assert r.err == 0, i
        assert hashlib.sha256(r.canonical.encode()).hexdigest() == row["nf_sha256"], i


def _ab(hip, ctx, run, seeds, w):
    """Plain and learned on alternating runs at `seeds`; run(ctx) evaluates
    and checks the result, returning the evaluation's ms."""
    ms = {hip.SCHED_PLAIN: [], hip.SCHED_LEARNED: []}
    for i, seed in enumerate(seeds):
        order = (hip.SCHED_PLAIN, hip.SCHED_LEARNED) if i % 2 == 0 else (hip.SCHED_LEARNED, hip.SCHED_PLAIN)
        for mode in order:
            learned = mode == hip.SCHED_LEARNED
            ctx.set_scheduler(mode, seed=seed, explore_pct=w.get("margin_pct", 0) if learned else 0,
                              params_fp16=w["params_fp16"] if learned else None)
            ms[mode].append(run(ctx))
    return float(np.median(ms[hip.SCHED_PLAIN])), float(np.median(ms[hip.SCHED_LEARNED]))


def test_trained_scheduler_gain(hip):
    """The committed scheduler weights (lambdavm-rust_amd/sched_weights.json:
    trained on church_loop(12, 12) at seeds 0-63 and on eight other workloads
    at seeds 0-7, each under every action held fixed) against the plain
    heuristic at held-out seeds 100-109, with identical normal form and ITRS
    in every run:
    * radix_lsd(18, 0), the LSD radix sort: learned at least 1% faster
      (measured 1.025x median over seeds 0-7, no seed below 1.018x: it takes
      the 16 times larger steady bound, worth 1.025x when held fixed, which
      lets the narrowing merges run their chains without turn ends);
    * church_loop(12, 12), the outer-loop Church workload: learned not slower
      than plain by more than 2% (measured 0.996x over seeds 0-7: with the
      flat-frontier rule plain is the best action held fixed there);
    * the 256-net batch: learned not slower than plain by more than 0.5%
      (no action held fixed beats plain by more than 0.3% there)."""
    path = os.path.join(ROOT, "lambdavm-rust_amd", "sched_weights.json")
    w = json.load(open(path))
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    case = next(c for c in g["workloads"] if c["generator"] == "church_loop" and c["params"] == {"n": 12, "m": 12})
    code, expect = P.church_loop(12, 12)

    def church(ctx):
        r = ctx.eval()
        assert r.itrs == case["itrs"] and r.result == str(expect)
        return r.eval_ms
    with hip.Context(time_limit_s=120) as ctx:
        ctx.load(hip.compile_book(code))
        ctx.eval()
        plain, learned = _ab(hip, ctx, church, range(100, 110), w)
    print(f"church_loop(12,12) median eval: plain {plain:.2f} ms, learned {learned:.2f} ms, "
          f"gain {plain / learned:.3f}x", flush=True)
    assert learned <= plain * 1.02

    case = next(c for c in g["workloads"] if c["generator"] == "radix_lsd" and c["params"] == {"k": 18, "seed": 0})
    code, expect = P.radix_lsd(18, 0)

    def lsd(ctx):
        r = ctx.eval()
        assert r.itrs == case["itrs"] and r.result == str(expect)
        return r.eval_ms
    with hip.Context(time_limit_s=120) as ctx:
        ctx.load(hip.compile_book(code))
        ctx.eval()
        plain, learned = _ab(hip, ctx, lsd, range(100, 110), w)
    print(f"radix_lsd(18,0) median eval: plain {plain:.2f} ms, learned {learned:.2f} ms, "
          f"gain {plain / learned:.3f}x", flush=True)
    assert learned <= plain * 0.99

    code, nets = P.batch(0, 256)
    fids = P.def_fids(code)
    entries = [fids[f"net{i}"] for i in range(len(nets))]
    est = [P.batch_estimate(gen, params) for gen, params, _, _ in nets]
    gb = json.load(open(os.path.join(ROOT, "tests", "golden", "batch.json")))

    def batch(ctx):
        per, tot = ctx.eval_batch(entries, est)
        assert tot.itrs == gb["total_itrs"] and tot.err == 0
        return tot.eval_ms
    with hip.Context(time_limit_s=300, node_cap=1 << 29, vars_cap=1 << 29, flags=hip.NO_READBACK) as ctx:
        ctx.load(hip.compile_book(code))
        batch(ctx)
        plain, learned = _ab(hip, ctx, batch, range(100, 110), w)
    print(f"batch(0, 256) median eval: plain {plain:.2f} ms, learned {learned:.2f} ms, "
          f"gain {plain / learned:.3f}x", flush=True)
    assert learned <= plain * 1.005


def test_autotuner_keeps_results(hip):
    """The autotuner's search on a small batch: every candidate reproduces the
    default's ITRS (it raises otherwise) and the chosen configuration is one
    of the measured ones."""
    import autotune as T
    w = T.batched(seed=5, n=12)
    rep = T.tune(w, repeats=1, log=print, base_cfg={},
                 space=[("waves", [512, 0]), ("idle_pct", [101, 0]), ("turn_steps", [96, 0]),
                        ("flags", [hip.COMPILED_CALLS, hip.LOCAL_NET, 0])])
    assert rep["itrs"] > 0
    assert rep["tuned_ms"] <= rep["default_ms"]
    assert any(m.get("cfg") == rep["config"] for m in rep["measurements"])
    # evaluator errors are raised, never recorded as a skipped candidate
    assert not any("flags" in m.get("error", "") for m in rep["measurements"])


# autotune.py workload name -> generator of the run-c records
TUNED = {"radix_lsd": "radix_lsd", "batch": None, "church_loop": "church_loop", "sum": "sum_tree",
         "bitonic": "bitonic", "radix": "radix"}


@pytest.mark.parametrize("name", sorted(TUNED))
def test_tuned_configurations_keep_parity(hip, name):
    """The autotuner's choice for each tuned workload (LSD radix sort of 2^18
    keys, the 256-net batch, church_loop(12, 12), sum tree d=20, bitonic sort
    of 2^16 keys, trie radix sort of 2^18 keys; lambdavm-rust_amd/
    tuned_<name>.json), compile-time variant included, evaluated in a child
    process that loads that build: ITRS and normal forms equal the run-c
    records."""
    import subprocess
    import sys
    path = os.path.join(ROOT, "lambdavm-rust_amd", f"tuned_{name}.json")
    if not os.path.exists(path):
        pytest.skip("no tuned configuration committed")
    cfg = json.load(open(path))["config"]
    env = dict(os.environ)
    if cfg.get("variant"):
        env["HVM_HIP_LIB"] = os.path.join(ROOT, "lambdavm-rust_amd", "lib", "variants",
                                          f"libhvm_hip_{cfg['variant']}.so")
    print("tuned", name, cfg, flush=True)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "lambdavm-rust_amd", "autotune.py"), "--child",
                        "--workload", name, "--cfg", json.dumps(cfg), "--repeats", "0", "--readback",
                        "--time-limit", "120"], capture_output=True, text=True, env=env, timeout=600)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert "error" not in res, res
    if cfg.get("variant"):
        assert res["lib"] == env["HVM_HIP_LIB"]
    if name == "batch":
        g = json.load(open(os.path.join(ROOT, "tests", "golden", "batch.json")))
        assert res["ident"][0] == g["total_itrs"]
        assert res["nets_nf_sha256"] == [row["nf_sha256"] for row in g["nets"]]
    else:
        g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
        case = next(c for c in g["workloads"] if c["generator"] == TUNED[name] and c.get("full_size"))
        assert res["ident"] == [case["itrs"], case["nf_sha256"]]
