"""The mid-turn redex pool (HVM_MID_POOL; the `pool` variant build,
lambdavm-rust_amd/lib/variants/libhvm_hip_pool.so).  Busy waves hand surplus
redexes to the output bag during a turn, and waves that run dry take them back
with an exchange with 0 before counting as idle.  It is off in the default
build (slower on every workload measured) and is kept parity-green here.  Each
case runs in a child process that loads the variant, and must reproduce the
ITRS and normal form of the run-c records.  The Church fan is the case that
first failed: an output-bag write left in the writer's L2 overwrote a slot
that a pull had already emptied, so the redex ran twice."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lambdavm-rust_amd", "lib", "variants", "libhvm_hip_pool.so")

# autotune.py workload name -> (golden generator, params)
CASES = {"church": ("church", {"n": 12, "m": 12}),
         "bitonic": ("bitonic", {"k": 16, "seed": 0}),
         "radix_lsd": ("radix_lsd", {"k": 18, "seed": 0})}

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CASES))
def test_mid_turn_pool_keeps_parity(hip, name):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is missing: build it with `make -C lambdavm-rust_amd variants`")
    gen, params = CASES[name]
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    case = next(c for c in g["workloads"] if c["generator"] == gen and c["params"] == params)
    env = dict(os.environ, HVM_HIP_LIB=LIB)
    print("pool", name, flush=True)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "lambdavm-rust_amd", "autotune.py"), "--child",
                        "--workload", name, "--cfg", json.dumps({"variant": "pool"}), "--repeats", "0",
                        "--readback", "--time-limit", "120"], capture_output=True, text=True, env=env, timeout=600)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert "error" not in res, res
    assert res["lib"] == LIB
    assert res["ident"] == [case["itrs"], case["nf_sha256"]]
