"""The reference-compatible C++ layer: callers of the reference API compile, link and run.

CPU: the reference's own example/simpleExample.cpp (when the reference tree is present in this
container) and the repository's examples compile and link against timeevolver_amd/cpp.
GPU: the repository's two-mode example reproduces the reference's golden CSV output byte for
byte through the C++ API, and a host observable's requestStopException gives status 3.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "timeevolver_amd", "cpp")
GOLD = os.path.join(ROOT, "tests", "golden")
REF_EXAMPLE = "/root/reference/example/simpleExample.cpp"


def _build_layer():
    r = subprocess.run(["make", "-s", "-C", CPP, "libtimeevolver.so", "examples/oscillators"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def test_layer_and_example_build():
    _build_layer()
    assert os.path.exists(os.path.join(CPP, "examples", "oscillators"))


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE), reason="reference tree not present")
def test_reference_simple_example_compiles_unchanged(tmp_path):
    _build_layer()
    exe = tmp_path / "simpleExample"
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(CPP, "include"), REF_EXAMPLE, "-o", str(exe),
                        "-L", CPP, "-ltimeevolver", "-Wl,-rpath," + CPP], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()


@pytest.mark.gpu
def test_oscillators_example_matches_golden_csv(tmp_path):
    exe = os.path.join(CPP, "examples", "oscillators")
    assert os.path.exists(exe), "build the C++ layer first"
    prefix = str(tmp_path / "osc_n")
    r = subprocess.run([exe, prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    for k in range(2):
        with open(prefix + "%d.csv" % k) as f, open(os.path.join(GOLD, "simple_example_mode%d.csv" % k)) as g:
            assert f.read() == g.read()
    stop = [l for l in r.stdout.splitlines() if l.startswith("stop:")][0]
    assert "status=3" in stop and "samples=5" in stop
