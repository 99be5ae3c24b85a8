"""Reads an HDF5 file through the h5dump tool's text output (h5py is not installed).

read(path) -> {dataset: {"type": str, "space": str, "data": [float], "attributes":
{name: {"type": str, "space": str, "data": [float]}}}}; dataset values are printed with 17
significant digits, so they round-trip to the stored doubles.
"""
import os
import re
import shutil
import subprocess

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def available():
    return os.path.exists(H5DUMP)


def _values(body):
    out = []
    for m in re.finditer(r"\(\d+\):\s*([^()]*)", body):
        out += [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    return out


def read(path):
    txt = subprocess.run([H5DUMP, "-m", "%.17g", path], check=True, capture_output=True, text=True).stdout
    out = {}
    ds = attr = None
    data_lines = None
    for raw in txt.splitlines():
        line = raw.strip()
        if data_lines is not None:
            if line == "}":
                tgt = out[ds] if attr is None else out[ds]["attributes"][attr]
                tgt["data"] = _values("\n".join(data_lines))
                data_lines = None
            else:
                data_lines.append(line)
            continue
        m = re.match(r'(DATASET|ATTRIBUTE) "([^"]+)"', line)
        if m and m.group(1) == "DATASET":
            ds, attr = m.group(2), None
            out[ds] = {"attributes": {}, "data": []}
        elif m:
            attr = m.group(2)
            out[ds]["attributes"][attr] = {"data": []}
        elif line.startswith("DATATYPE") or line.startswith("DATASPACE"):
            key = "type" if line.startswith("DATATYPE") else "space"
            (out[ds] if attr is None else out[ds]["attributes"][attr])[key] = line.split(None, 1)[1]
        elif line.startswith("DATA {"):
            data_lines = []
        elif line == "}" and attr is not None:
            attr = None
    return out
