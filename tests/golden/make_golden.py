"""Writes the golden fixtures under tests/golden/ from the reference's own output files.

Run in a container that has the reference mounted at /root/reference and h5dump at
/opt/conda/bin/h5dump. The fixtures are data only (inputs are the reference's example
parameters, outputs the observables its ctests assert at 1e-8):

  blackhole_N20_Nm2_K4.txt  <- output/ResultBlackHole_N20_..._fastIntegration0.h5
                               (test/testBlackHole.cpp:21-45), 10 modes x 1001 samples
  blackhole_N1_Nm1_K1.txt   <- output/ResultBlackHole_N1_..._fastIntegration0.h5
                               (test/testBlackHoleLB.cpp:22-46), 4 modes x 1001 samples
  simple_example_mode{0,1}.csv <- output/SimpleExampleOutputOccupationNumber{0,1}.csv
                               (test/testSimpleExample.cpp:36-76), copied byte for byte
  result_csv/<name>         <- output/<name> for three of the CSV files the reference's
                               saveResult wrote without HDF5 (helper/krylovHelper.cpp:94-104),
                               copied byte for byte
  blackhole_N20_Nm2_K4_h5_layout.json <- h5dump -H of the N20 .h5: each dataset's type and
                               shape, each attribute's type and shape
Each .txt line is one dataset "modeK", values printed with 17 significant digits.
"""
import json
import os
import re
import shutil
import subprocess

REF = "/root/reference/output"
OUT = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"

FILES = {
    "blackhole_N20_Nm2_K4.txt": (
        "ResultBlackHole_N20_Nm2_K4_C1_DeltaN12_C01_Cm1_maxT10_tol1e-08_samplingStep0.01_m40_fastIntegration0.h5", 10),
    "blackhole_N1_Nm1_K1.txt": (
        "ResultBlackHole_N1_Nm1_K1_C1_DeltaN12_C01_Cm1_maxT1000_tol1e-08_samplingStep1_m40_fastIntegration0.h5", 4),
}


def dump(path, ds):
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-d", "/" + ds, path], check=True, capture_output=True,
                         text=True).stdout
    body = txt.split("DATA {", 1)[1].split("}", 1)[0]
    vals = re.findall(r"\(\d+\):\s*([^\n]+)", body)
    out = []
    for line in vals:
        out += [v.strip() for v in line.split(",") if v.strip()]
    return out


CSV_FILES = [
    "ResultBlackHole_N20_Nm2_K4_C1_DeltaN12_C01_Cm1_maxT10_tol1e-08_samplingStep0.01_m40_fastIntegration0mode0.csv",
    "ResultBlackHole_N20_Nm2_K4_C1_DeltaN12_C01_Cm1_maxT10_tol1e-08_samplingStep0.01_m40_fastIntegration0mode5.csv",
    "ResultBlackHole_N1_Nm1_K1_C1_DeltaN12_C01_Cm1_maxT1000_tol1e-08_samplingStep1_m40_fastIntegration0mode1.csv",
]


def layout(path):
    """{dataset: {"type", "space", "attributes": {name: [type, space]}}} from h5dump -H."""
    txt = subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout
    out, ds, attr = {}, None, None
    for line in txt.splitlines():
        line = line.strip()
        m = re.match(r'(DATASET|ATTRIBUTE) "([^"]+)"', line)
        if m and m.group(1) == "DATASET":
            ds, attr = m.group(2), None
            out[ds] = {"attributes": {}}
        elif m:
            attr = m.group(2)
            out[ds]["attributes"][attr] = [None, None]
        elif line.startswith("DATATYPE") or line.startswith("DATASPACE"):
            key = 0 if line.startswith("DATATYPE") else 1
            val = line.split(None, 1)[1]
            if attr is None:
                out[ds]["type" if key == 0 else "space"] = val
            else:
                out[ds]["attributes"][attr][key] = val
        elif line == "}" and attr is not None:
            attr = None
    return out


def main():
    for name, (h5, nmodes) in FILES.items():
        lines = [" ".join(dump(os.path.join(REF, h5), "mode%d" % k)) for k in range(nmodes)]
        with open(os.path.join(OUT, name), "w") as f:
            f.write("\n".join(lines) + "\n")
    for k in (0, 1):
        shutil.copyfile(os.path.join(REF, "SimpleExampleOutputOccupationNumber%d.csv" % k),
                        os.path.join(OUT, "simple_example_mode%d.csv" % k))
    os.makedirs(os.path.join(OUT, "result_csv"), exist_ok=True)
    for name in CSV_FILES:
        shutil.copyfile(os.path.join(REF, name), os.path.join(OUT, "result_csv", name))
    with open(os.path.join(OUT, "blackhole_N20_Nm2_K4_h5_layout.json"), "w") as f:
        json.dump(layout(os.path.join(REF, FILES["blackhole_N20_Nm2_K4.txt"][0])), f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
