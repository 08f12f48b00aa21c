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
with canonical var names, and ITRS.
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
# Files the programs open, by name, with contents of this repository's own.
FILES = {
    "LICENSE": "HVM IO fixture: these forty-seven bytes are read by the IO tests.\nsecond line\n",
    "README.md": "fixture\n",
}


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
        print(name, rec["result"], rec["itrs"], rec["lines"])
        out["programs"].append(rec)
    with open(os.path.join(HERE, "io.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
