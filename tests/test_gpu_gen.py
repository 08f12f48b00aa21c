"""`hvm gen-hip` output run on the GPU.

build() emits the standalone program of the reference's
examples/sort_bitonic/main.hvm with `hvm gen-hip` (the book as a static array
and a main that calls the drop-in entry `hvm_hip`, the analogue of gen-cu) and
compiles it against lib/libhvm_hip.so.  Run here, its stdout must carry the
reference run-c's result (canonical var names) and ITRS
(tests/golden/golden.json) in run-cu's format.
"""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lambdavm-rust_amd", "bin", "gen_sort_bitonic")


def test_gen_hip_program_matches_run_c(hip):
    if not os.access(EXE, os.X_OK):
        pytest.fail("bin/gen_sort_bitonic is not built (build() makes it)")
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    want = next(p for p in g["programs"] if p["name"] == "examples/sort_bitonic/main.hvm")
    print("gen-hip program: examples/sort_bitonic/main.hvm", flush=True)
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = p.stdout.splitlines()
    result = next(line[len("Result: "):] for line in lines if line.startswith("Result: "))
    itrs = next(int(line[len("- ITRS: "):]) for line in lines if line.startswith("- ITRS: "))
    assert [line.split(":")[0] for line in lines[1:]] == ["- ITRS", "- LEAK", "- TIME", "- MIPS"]
    assert itrs == want["itrs"]
    assert hashlib.sha256(hip.canonical(result).encode()).hexdigest() == want["nf_sha256"]
