"""CPU checks of the committed artefacts the GPU tests and the bench load:
the scheduler weights (lambdavm-rust_amd/sched_weights.json) and the
autotuner's configurations (lambdavm-rust_amd/tuned_*.json).  A malformed
file would otherwise only show up as a failure on the GPU."""
import glob
import json
import os
import re

import numpy as np

import mlp_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lambdavm-rust_amd")


def header_define(name):
    text = open(os.path.join(ROOT, "include", "hvm_hip.h")).read()
    m = re.search(rf"#define {name}\s+(.+)", text)
    return m.group(1).strip()


def test_scheduler_weights_fit_the_network():
    w = json.load(open(os.path.join(PKG, "sched_weights.json")))
    p = w["params_fp16"]
    # HVM_HIP_MLP_PARAMS = (32 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4)
    assert len(p) == eval(header_define("HVM_HIP_MLP_PARAMS"), {}) == sum(mlp_ref.SIZES)
    assert all(isinstance(v, int) and 0 <= v < 1 << 16 for v in p)
    vals = np.array(p, dtype=np.uint16).view(np.float16).astype(np.float32)
    assert np.isfinite(vals).all() and np.abs(vals).max() < 1e3
    assert isinstance(w["margin_pct"], int) and 0 <= w["margin_pct"] <= 100
    rep = w["report"]
    assert rep["identical_results"] is True
    assert all(row["identical_results"] for row in rep["per_workload"].values())


def test_tuned_configurations_name_known_knobs():
    knobs = {"waves", "idle_pct", "turn_steps", "flags", "graphs", "variant", "node_cap", "vars_cap"}
    makefile = open(os.path.join(PKG, "Makefile")).read()
    variants = set(re.findall(r"(\w+):-D", makefile))
    files = sorted(glob.glob(os.path.join(PKG, "tuned_*.json")))
    assert {os.path.basename(f) for f in files} >= {"tuned_radix_lsd.json", "tuned_batch.json",
                                                    "tuned_church_loop.json"}
    for f in files:
        rep = json.load(open(f))
        cfg = rep["config"]
        assert set(cfg) <= knobs, (f, cfg)
        assert not cfg.get("variant") or cfg["variant"] in variants, (f, cfg)
        assert 0 < rep["tuned_ms"] <= rep["default_ms"], f
        # every candidate the search kept reproduced the workload's results
        assert any(m.get("cfg") == cfg for m in rep["measurements"]), f
