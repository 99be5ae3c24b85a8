"""The reference-compatible C++ layer: callers of the reference API compile, link and run.

CPU: the reference's own example/simpleExample.cpp (when the reference tree is present in this
container) and the repository's examples compile and link against timeevolver_amd/cpp.
Result output (saveResult, helper/krylovHelper.cpp:10-106), from host data: the CSV build
rewrites the reference's own CSV result files byte for byte (file names included) from the
values of its .h5 results; the HDF5 build writes the reference's .h5 file name, dataset types
and shapes, and an attribute of the reference's type per int/bool/double parameter.
GPU: the repository's two-mode example reproduces the reference's golden CSV output byte for
byte through the C++ API, and a host observable's requestStopException gives status 3; the
black-hole example (example/main.cpp's flow) writes the reference's .h5 result within 1e-8.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import h5text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "timeevolver_amd", "cpp")
GOLD = os.path.join(ROOT, "tests", "golden")
REF_EXAMPLE = "/root/reference/example/simpleExample.cpp"


def _build_layer():
    r = subprocess.run(["make", "-s", "-C", CPP, "libtimeevolver.so", "examples/oscillators", "examples/blackhole"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


HDF5_DIR = "/opt/conda"
HAVE_HDF5 = os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")) and h5text.available()

# example/main.cpp's parameters for the two reference result files, in its paraPush order
BH_RUNS = {
    "blackhole_N20_Nm2_K4.txt": ["20", "2", "4", "1", "12", "1", "1", "10", "1e-8", "0.01", "40", "0"],
    "blackhole_N1_Nm1_K1.txt": ["1", "1", "1", "1", "12", "1", "1", "1000", "1e-8", "1", "40", "0"],
}
BH_STEM = {
    "blackhole_N20_Nm2_K4.txt":
        "ResultBlackHole_N20_Nm2_K4_C1_DeltaN12_C01_Cm1_maxT10_tol1e-08_samplingStep0.01_m40_fastIntegration0",
    "blackhole_N1_Nm1_K1.txt":
        "ResultBlackHole_N1_Nm1_K1_C1_DeltaN12_C01_Cm1_maxT1000_tol1e-08_samplingStep1_m40_fastIntegration0",
}
# attribute names of the reference's stored .h5 (an earlier parameter naming) -> example/main.cpp's
REF_ATTR = {"N0": "N", "Capacity": "C", "MaxT": "maxT", "SamplingStep": "samplingStep", "Tolerance": "tol"}


def _build_writer(tmp_path, hdf5):
    """tests/cpp/save_output.cpp. HDF5 flavour: linked against the layer as the Makefile builds it
    (with HDF5 when the library is found); CSV flavour: the whole layer compiled in without it."""
    _build_layer()
    exe = str(tmp_path / ("save_output_h5" if hdf5 else "save_output_csv"))
    src = [os.path.join(ROOT, "tests", "cpp", "save_output.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(CPP, "include")]
    if hdf5:
        cmd += src + ["-o", exe, "-L", CPP, "-ltimeevolver", "-Wl,-rpath," + CPP,
                      "-Wl,-rpath," + os.path.join(ROOT, "timeevolver_amd")]
    else:
        src += sorted(os.path.join(CPP, "src", f) for f in os.listdir(os.path.join(CPP, "src")) if f.endswith(".cpp"))
        cmd += src + ["-o", exe, "-L", os.path.join(ROOT, "timeevolver_amd"), "-ltimeevolver_hip",
                      "-Wl,-rpath," + os.path.join(ROOT, "timeevolver_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_result_csv_matches_reference_files(tmp_path):
    exe = _build_writer(tmp_path, hdf5=False)
    for golden, args in BH_RUNS.items():
        out = tmp_path / golden.split(".")[0]
        out.mkdir()
        _run([exe, str(out), "result", os.path.join(GOLD, golden)] + args)
        n_modes = len(open(os.path.join(GOLD, golden)).read().splitlines())
        assert sorted(os.listdir(out)) == sorted(BH_STEM[golden] + "mode%d.csv" % k for k in range(n_modes))
    checked = 0
    for name in os.listdir(os.path.join(GOLD, "result_csv")):
        sub = "blackhole_N20_Nm2_K4" if name.startswith(BH_STEM["blackhole_N20_Nm2_K4.txt"]) else "blackhole_N1_Nm1_K1"
        with open(tmp_path / sub / name, "rb") as f, open(os.path.join(GOLD, "result_csv", name), "rb") as g:
            assert f.read() == g.read(), name
        checked += 1
    assert checked == 3


def test_generic_output_csv_build(tmp_path):
    exe = _build_writer(tmp_path, hdf5=False)
    out = _run([exe, str(tmp_path), "generic"])
    assert "no HDF5 support" in out
    stem = tmp_path / "Res_N20_tol1e-08_C01_fast0_u7_labelx_pi3.141592654"
    assert open(str(stem) + "alpha.csv").read() == "1, -2.5, 1e-300, 0.1, 20"
    assert open(str(stem) + "beta.csv").read() == "3"


def _check_reference_layout(ours, ref):
    """Dataset names, types and shapes as in the reference file; each reference attribute present
    (under example/main.cpp's name) with the reference's type and shape."""
    assert sorted(ours) == sorted(ref)
    for ds, d in ref.items():
        assert ours[ds]["type"] == d["type"] and ours[ds]["space"] == d["space"], ds
        for name, (typ, space) in d["attributes"].items():
            a = ours[ds]["attributes"][REF_ATTR.get(name, name)]
            assert (a["type"], a["space"]) == (typ, space), (ds, name)


@pytest.mark.parametrize("hdf5", [False, True])
def test_writers_clean_under_sanitizers(tmp_path, hdf5):
    """The result writers (both flavours) with AddressSanitizer, leak detection and
    UndefinedBehaviorSanitizer: host data only, so the whole layer runs on the CPU."""
    if hdf5 and not HAVE_HDF5:
        pytest.skip("HDF5 library or h5dump not present")
    _build_layer()
    exe = str(tmp_path / "save_output_san")
    src = [os.path.join(ROOT, "tests", "cpp", "save_output.cpp")]
    src += sorted(os.path.join(CPP, "src", f) for f in os.listdir(os.path.join(CPP, "src")) if f.endswith(".cpp"))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", os.path.join(CPP, "include")] + src + ["-o", exe,
           "-L", os.path.join(ROOT, "timeevolver_amd"), "-ltimeevolver_hip",
           "-Wl,-rpath," + os.path.join(ROOT, "timeevolver_amd")]
    if hdf5:
        # the HDF5 prefix also holds an older libstdc++: reach libhdf5 through a directory of its own
        libdir = tmp_path / "h5lib"
        libdir.mkdir()
        real = os.path.realpath(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))
        os.symlink(real, libdir / os.path.basename(real))
        soname = [f for f in os.listdir(os.path.join(HDF5_DIR, "lib")) if f.startswith("libhdf5.so.")
                  and f.count(".") == 2]
        for name in soname:
            if not (libdir / name).exists():
                os.symlink(real, libdir / name)
        cmd += ["-DUSE_HDF", "-idirafter", HDF5_DIR + "/include", os.path.join(HDF5_DIR, "lib", "libhdf5.so"),
                "-Wl,-rpath," + str(libdir)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = tmp_path / "out"
    out.mkdir()
    for args in (["generic"], ["result", os.path.join(GOLD, "blackhole_N20_Nm2_K4.txt")] +
                 BH_RUNS["blackhole_N20_Nm2_K4.txt"]):
        r = subprocess.run([exe, str(out)] + args, capture_output=True, text=True, env=env, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr


@pytest.mark.skipif(not HAVE_HDF5, reason="HDF5 library or h5dump not present")
def test_result_h5_matches_reference_layout(tmp_path):
    exe = _build_writer(tmp_path, hdf5=True)
    golden = "blackhole_N20_Nm2_K4.txt"
    _run([exe, str(tmp_path), "result", os.path.join(GOLD, golden)] + BH_RUNS[golden])
    assert os.listdir(tmp_path / ".") and (tmp_path / (BH_STEM[golden] + ".h5")).exists()
    ours = h5text.read(str(tmp_path / (BH_STEM[golden] + ".h5")))
    with open(os.path.join(GOLD, "blackhole_N20_Nm2_K4_h5_layout.json")) as f:
        ref = {k: {"type": v["type"], "space": v["space"], "attributes": v["attributes"]} for k, v in json.load(f).items()}
    _check_reference_layout(ours, ref)
    g = np.loadtxt(os.path.join(GOLD, golden))
    for k in range(len(g)):
        assert np.array_equal(np.array(ours["mode%d" % k]["data"]), g[k])
    a = ours["mode0"]["attributes"]
    assert [a[n]["data"][0] for n in ("N", "Nm", "K", "C", "m", "fastIntegration")] == [20, 2, 4, 1, 40, 0]
    assert [a[n]["data"][0] for n in ("DeltaN", "C0", "Cm", "maxT", "tol", "samplingStep")] == [12, 1, 1, 10, 1e-8, 0.01]


@pytest.mark.skipif(not HAVE_HDF5, reason="HDF5 library or h5dump not present")
def test_generic_output_h5_build(tmp_path):
    exe = _build_writer(tmp_path, hdf5=True)
    assert "matrices: written" in _run([exe, str(tmp_path), "generic"])
    res = h5text.read(str(tmp_path / "Res_N20_tol1e-08_C01_fast0_u7_labelx_pi3.141592654.h5"))
    assert sorted(res) == ["alpha", "beta"]
    assert res["alpha"]["data"] == [1.0, -2.5, 1e-300, 0.1, 20.000000000000004]
    assert res["beta"]["data"] == [3.0] and res["beta"]["space"] == "SIMPLE { ( 1 ) / ( 1 ) }"
    for ds in res.values():
        at = ds["attributes"]
        # unsigned and string parameters are named in the file but get no attribute
        assert sorted(at) == ["C0", "N", "dt", "fast", "pi", "tol"]
        assert {n: a["type"] for n, a in at.items()} == {
            "C0": "H5T_IEEE_F64LE", "N": "H5T_STD_I32LE", "dt": "H5T_IEEE_F64LE", "fast": "H5T_STD_I32LE",
            "pi": "H5T_IEEE_F64LE", "tol": "H5T_IEEE_F64LE"}
        assert [at[n]["data"][0] for n in ("N", "fast", "dt", "tol", "C0")] == [20, 0, 0.01, 1e-8, 1.0]
    sp = h5text.read(str(tmp_path / "sparse.h5"))
    assert sp["valuesRealPart"]["data"] == [1, 3, 0.5, 0] and sp["valuesImagPart"]["data"] == [2, -4, 0, -1e-3]
    assert sp["columIndex"]["data"] == [0, 2, 1, 2] and sp["rowIndex"]["data"] == [0, 0, 2, 1]
    assert sp["dimension"]["data"] == [3] and sp["rowIndex"]["type"] == "H5T_STD_U64LE"
    de = h5text.read(str(tmp_path / "dense.h5"))
    assert de["valuesRealPart"]["data"] == [1, 3, 0.5, 0] and de["valuesImagPart"]["data"] == [2, -4, 0, -1e-3]


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


@pytest.mark.gpu
@pytest.mark.skipif(not HAVE_HDF5, reason="HDF5 library or h5dump not present")
def test_blackhole_example_writes_reference_h5(tmp_path):
    """example/main.cpp's run at its defaults: the .h5 carries the reference's file name and layout
    and its values within the reference test's 1e-8 (test/testBlackHole.cpp)."""
    exe = os.path.join(CPP, "examples", "blackhole")
    assert os.path.exists(exe), "build the C++ layer first"
    out = _run([exe, "--out", str(tmp_path / "ResultBlackHole")])
    print(out)
    golden = "blackhole_N20_Nm2_K4.txt"
    assert os.listdir(tmp_path) == [BH_STEM[golden] + ".h5"]
    ours = h5text.read(str(tmp_path / (BH_STEM[golden] + ".h5")))
    with open(os.path.join(GOLD, "blackhole_N20_Nm2_K4_h5_layout.json")) as f:
        ref = json.load(f)
    _check_reference_layout(ours, ref)
    g = np.loadtxt(os.path.join(GOLD, golden))
    diff = max(np.abs(np.array(ours["mode%d" % k]["data"]) - g[k]).max() for k in range(len(g)))
    print("max |diff| vs reference %.3e" % diff)
    assert diff <= 1e-8
    a = ours["mode0"]["attributes"]
    assert a["TE_MAJOR_VERSION"]["data"] == [2] and a["TE_MINOR_VERSION"]["data"] == [1]
