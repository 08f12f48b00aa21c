"""Resource use of the HBM turn kernels, read from their gfx950 code objects.

The turn kernels are latency-bound: a scratch reload issued behind a
write-through store waits for that store (the vector-memory counter counts
loads and stores in issue order), so a register spill on the step's path costs
a memory round trip, and the kernels are sized for 8 resident waves per CU
(2 per SIMD: at most 256 VGPRs, at most 160 KiB / 8 of LDS each).  This test
compiles hvm_eval.hip for gfx950 (no GPU needed) and checks those limits from
the assembler's kernel metadata.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "lambdavm-rust_amd", "csrc", "hip", "hvm_eval.hip")
HIPCC = "/opt/rocm/bin/hipcc"

KERNELS = ("_Z15hvm_turn_kernel11HvmTurnArgs", "_Z20hvm_turn_fast_kernel11HvmTurnArgs")


def kernel_metadata(asm: str):
    """Per kernel: the metadata fields of the AMDGPU assembler's kernel list."""
    out = {}
    for block in asm.split("  - .agpr_count:")[1:]:
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue
        fields = dict(re.findall(r"\.(\w+):\s+(\d+)\n", block))
        out[m.group(1)] = {k: int(v) for k, v in fields.items()}
    return out


@pytest.fixture(scope="module")
def meta(tmp_path_factory):
    if not os.access(HIPCC, os.X_OK) and not shutil.which("hipcc"):
        pytest.skip("hipcc is not installed")
    out = tmp_path_factory.mktemp("isa") / "hvm_eval.s"
    cmd = [HIPCC if os.access(HIPCC, os.X_OK) else "hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "--cuda-device-only", "-S", SRC, "-o", str(out)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    return kernel_metadata(out.read_text())


@pytest.mark.parametrize("kernel", KERNELS)
def test_turn_kernel_has_no_spills(meta, kernel):
    m = meta[kernel]
    assert m["vgpr_spill_count"] == 0, m
    assert m["private_segment_fixed_size"] == 0, m


@pytest.mark.parametrize("kernel", KERNELS)
def test_turn_kernel_fits_eight_waves_per_cu(meta, kernel):
    m = meta[kernel]
    assert m["vgpr_count"] <= 256, m
    assert m["group_segment_fixed_size"] * 8 <= 160 * 1024, m
