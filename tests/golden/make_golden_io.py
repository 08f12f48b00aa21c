"""Writes tests/golden/io.json: the reference run-c's output on IO programs.

    python tests/golden/make_golden_io.py   (needs oracle/_ref, built by `make -C oracle`
                                             where /root/reference exists)

Programs: the reference's IO tests (tests/programs/io/*.hvm and
examples/demo_io/main.hvm of the reference, read from /root/reference) and
tests/programs/io/echo_seek.hvm of this repository (READ from stdin, WRITE,
FLUSH, SEEK on a bad descriptor, an unknown IO function).  Each runs in a
scratch directory holding the files of FILES, with STDIN on standard input,
as the reference's tests/run.rs:67-110 runs them.  Recorded per program: the
program text, the output lines other than ITRS/TIME/MIPS/LEAK, the result
with canonical var names, ITRS, and the reference's own snapshot of the
program's output where it has one.
"""
import glob
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lambdavm-rust_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hvm_hip as H  # noqa: E402
import oracle  # noqa: E402

REF = "/root/reference"
STDIN = "io from the tests\n"
# Files the programs open, by name.  LICENSE starts as the Apache License 2.0
# text does (the file the reference's IO tests read from its repository root),
# so that run-c's output can also be checked against the reference's own
# snapshots of these programs (tests/snapshots/run__io_file@*.snap).
FILES = {
    "LICENSE": "                                 Apache License\n                           Version 2.0, January 2004\n",
    "README.md": "fixture\n",
}


def reference_snapshot(name):
    """The normalised output the reference's tests/run.rs recorded for an IO
    program (lines after the snapshot header; None if there is none)."""
    path = os.path.join(REF, "tests/snapshots", "run__io_file@" + name.replace("tests/programs/", "").replace("examples/", "")
                        .replace("/", "__") + ".snap")
    if not os.path.exists(path):
        return None
    text = open(path).read().split("---\n", 2)[2]
    lines = text.rstrip("\n").split("\n")
    result = [l[len("Result: "):] for l in lines if l.startswith("Result: ")]
    return {"lines": [l for l in lines if not l.startswith("Result: ")],
            "result": H.canonical(result[0]) if result else None}


def run_reference(code):
    words = H.compile_book(code)
    with tempfile.TemporaryDirectory() as d:
        for name, text in FILES.items():
            with open(os.path.join(d, name), "w") as f:
                f.write(text)
        book = os.path.join(d, "prog.book")
        with open(book, "wb") as f:
            f.write(b"".join(int(w).to_bytes(4, "little") for w in words))
        p = subprocess.run([oracle.ref_binary(1), book], input=STDIN, capture_output=True, text=True,
                           cwd=d, timeout=120)
    lines, result, itrs = [], None, None
    for line in (p.stdout + p.stderr).splitlines():
        if line.startswith("Result: "):
            result = H.canonical(line[len("Result: "):])
        elif line.startswith("- ITRS: "):
            itrs = int(line[len("- ITRS: "):])
        elif not line.startswith(("- TIME:", "- MIPS:", "- LEAK:")):
            lines.append(line)
    return {"lines": lines, "result": result, "itrs": itrs, "returncode": p.returncode}


def main():
    progs = sorted(glob.glob(os.path.join(REF, "tests/programs/io/*.hvm"))) + \
        [os.path.join(REF, "examples/demo_io/main.hvm"), os.path.join(ROOT, "tests/programs/io/echo_seek.hvm")]
    out = {"stdin": STDIN, "files": FILES, "programs": []}
    for path in progs:
        code = open(path).read()
        name = os.path.relpath(path, REF) if path.startswith(REF) else os.path.relpath(path, ROOT)
        rec = dict(name=name, code=code, **run_reference(code))
        snap = reference_snapshot(name) if path.startswith(REF) else None
        if snap is not None:
            rec["reference_snapshot"] = snap
        print(name, rec["result"], rec["itrs"], rec["lines"])
        out["programs"].append(rec)
    with open(os.path.join(HERE, "io.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
