"""The C-ABI library: loads, exports every entry point include/hvm_hip.h
declares, and the CLI host builds book buffers identical to the Python path."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT, PKG


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hvm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hvm_[a-z_]+)\s*\(", text)))


def test_header_declares_the_drop_in():
    names = declared_functions()
    assert "hvm_hip" in names and "hvm_hip_run" in names
    assert len(names) >= 9


def test_library_exports_every_declared_symbol(hip):
    L = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing


def test_cli_compile_matches_library(hip, tmp_path):
    exe = os.path.join(PKG, "bin", "hvm")
    if not os.access(exe, os.X_OK):
        pytest.skip("CLI not built")
    src = tmp_path / "p.hvm"
    src.write_text("@main = r & @f ~ (3 r)\n@f = ({a b} c) & a ~ $([+] $(b c))\n")
    out = subprocess.run([exe, "compile", str(src)], capture_output=True, check=True).stdout
    words = [int.from_bytes(out[i:i + 4], "little") for i in range(0, len(out), 4)]
    assert words == hip.compile_book(src.read_text())


def test_gen_hip_emits_a_buildable_program(hip, tmp_path):
    exe = os.path.join(PKG, "bin", "hvm")
    if not os.access(exe, os.X_OK):
        pytest.skip("CLI not built")
    src = tmp_path / "p.hvm"
    src.write_text("@main = r & @f ~ (3 r)\n@f = ({a b} c) & a ~ $([+] $(b c))\n")
    out = subprocess.run([exe, "gen-hip", str(src)], capture_output=True, text=True, check=True).stdout
    body = out.split("{", 1)[1].split("}", 1)[0]
    words = [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})", body)]
    assert words == hip.compile_book(src.read_text())
    cpp = tmp_path / "gen.cpp"
    cpp.write_text(out)
    subprocess.run(["g++", str(cpp), "-I", os.path.join(ROOT, "include"), "-L", os.path.join(PKG, "lib"),
                    "-lhvm_hip", "-o", str(tmp_path / "gen")], check=True)


def test_cli_reports_missing_main(tmp_path):
    exe = os.path.join(PKG, "bin", "hvm")
    if not os.access(exe, os.X_OK):
        pytest.skip("CLI not built")
    src = tmp_path / "empty.hvm"
    src.write_text("")
    p = subprocess.run([exe, "run-hip", str(src)], capture_output=True, text=True)
    assert p.returncode == 101
    assert "missing `@main` definition" in p.stderr


def test_no_device_is_an_error_not_a_fallback(hip):
    import ctypes as C
    L = hip.lib()
    # Without a GPU the context cannot be created and the library says why;
    # there is no CPU path behind it.
    has_gpu = os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)
    if has_gpu:
        pytest.skip("a GPU is present")
    h = L.hvm_hip_create(None)
    assert not h
    assert b"device" in L.hvm_hip_last_error(None).lower()
    with pytest.raises(hip.HvmError):
        hip.run_code("@main = 1")


def test_bench_refuses_more_than_one_gpu():
    # One GPU per box: --gpus above 1 exits with a message before any GPU work.
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_wide_mode_is_an_error_without_a_device(hip):
    has_gpu = os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(hip.HvmError):
        hip.Context(flags=hip.WIDE)
