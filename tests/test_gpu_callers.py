"""The reference's own callers of qdt.h, compiled against include/qdt.h and
linked to libqdt_amd.so (oracle/Makefile, target `callers`): its example
program (example/example.c) and its test program (test/qdt_verify.cpp),
which checks every result with exact integer arithmetic and has every
scenario of the API (batch, guards, insertion, constraints, contexts).
Exit status 0 is its verdict "all checks passed"."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.gpu


def need(name):
    p = os.path.join(REF, name)
    if not os.path.exists(p):
        pytest.skip("%s not built (needs the reference sources at build time)" % name)
    return p


def run(args, timeout):
    print("run:", " ".join(os.path.basename(a) for a in args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(r.stderr[-3000:], flush=True)
    return r


def canon(tris):
    out = []
    for t in tris:
        k = t.index(min(t))
        out.append(tuple(t[k:] + t[:k]))
    return sorted(out)


def test_example_program(gpu):
    r = run([need("example_amd")], 120)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[:3] == ["Point 4 inserted", "Point 5 inserted", "Triangles: 6"]
    tris = [[int(x) for x in l.split(":")[1].split(",")] for l in lines[3:]]
    # the reference README's listing of this program's output (README.md:473-482)
    want = [[1, 3, 4], [0, 1, 4], [2, 0, 5], [0, 4, 5], [4, 3, 5], [3, 2, 5]]
    assert canon(tris) == canon(want)


def test_qdt_verify_all_scenarios_quick(gpu):
    r = run([need("qdt_verify_amd"), "--quick"], 900)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "all checks passed" in r.stderr


def test_qdt_verify_big(gpu):
    r = run([need("qdt_verify_amd"), "--scenario", "big,contexts"], 900)
    assert r.returncode == 0, r.stdout[-2000:]
