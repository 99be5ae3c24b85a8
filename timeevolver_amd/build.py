"""Builds libtimeevolver_hip.so (gfx950) in-tree with hipcc, and the C++ drop-in layer.

    python -m timeevolver_amd.build          # or __graft_entry__.build()

Every HIP translation unit is compiled with packed-FP32 instructions disabled. The step
controller and the matrix builder are compiled with -ffp-contract=off: their scalar
arithmetic must round exactly as the reference's x86-64 build (no FMA) does.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtimeevolver_hip.so")
ARCH = os.environ.get("TE_OFFLOAD_ARCH", "gfx950")

SOURCES = {
    "te_capi.hip": [],
    "te_spmv.hip": [],
    "te_krylov.hip": ["-ffp-contract=off"],
    "te_builder.hip": ["-ffp-contract=off"],
    "te_batch.hip": [],
    "te_adjoint.hip": [],
}

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include")]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def _compile(src, extra, outdir):
    obj = os.path.join(outdir, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build_hip(verbose=False):
    outdir = os.path.join(HERE, "_build")
    os.makedirs(outdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, SOURCES[s], outdir), srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cpp(verbose=False):
    """The reference-compatible C++ layer (cpp/) and its examples; result files are HDF5 when
    the Makefile finds the HDF5 library, CSV otherwise."""
    mk = os.path.join(HERE, "cpp", "Makefile")
    if not os.path.exists(mk):
        return None
    r = subprocess.run(["make", "-s", "-C", os.path.dirname(mk), "libtimeevolver.so", "examples/oscillators",
                        "examples/blackhole"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ layer build failed:\n%s%s" % (r.stdout, r.stderr))
    return os.path.join(HERE, "cpp", "libtimeevolver.so")


def build(verbose=False):
    build_hip(verbose)
    build_cpp(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
