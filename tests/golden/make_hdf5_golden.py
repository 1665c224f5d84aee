"""Expected contents of the reference's HDF5 fixture, as read by h5dump (the
HDF Group's own reader, independent of itg.hdf5).

tests/golden/hdf5/testfilev0.1.41.h5 is a copy of the reference's data file
test/base/utils/testfilev0.1.41.h5 (an ITensor written by ITensors.jl
v0.1.41: a Dense{Float64} 2x3x4 tensor under the old "store" key, read by
the reference's test/base/test_readwrite.jl:182-192). This script writes
tests/golden/hdf5/testfilev0.1.41.json: index ids, dims, dirs, plevs, tags
and the 24 data values, parsed from h5dump's text output.

usage: python tests/golden/make_hdf5_golden.py [path/to/h5dump]
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "hdf5", "testfilev0.1.41.h5")
TOP = "/ITensorv0.1.41"


def dump(h5dump, path):
    """The values of one dataset as h5dump prints them (floats with 17
    significant digits), as strings."""
    out = subprocess.run([h5dump, "-m", "%.17g", "-d", path, H5], check=True,
                         capture_output=True, text=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    vals = re.findall(r'"[^"]*"|[-+0-9.eE]+(?=[,\s])', re.sub(r"\(\d+\):", " ", body) + " ")
    return [v.strip('"') for v in vals]


def main():
    h5dump = sys.argv[1] if len(sys.argv) > 1 else "/opt/conda/bin/h5dump"
    n = int(dump(h5dump, f"{TOP}/inds/length")[0])
    inds = []
    for k in range(1, n + 1):
        g = f"{TOP}/inds/index_{k}"
        inds.append({"id": str(int(dump(h5dump, f"{g}/id")[0])), "dim": int(dump(h5dump, f"{g}/dim")[0]),
                     "dir": int(dump(h5dump, f"{g}/dir")[0]), "plev": int(dump(h5dump, f"{g}/plev")[0]),
                     "tags": dump(h5dump, f"{g}/tags/tags")[0]})
    data = [float(v) for v in dump(h5dump, f"{TOP}/store/data")]
    with open(os.path.join(HERE, "hdf5", "testfilev0.1.41.json"), "w") as f:
        json.dump({"group": TOP[1:], "storage_key": "store", "type": "Dense{Float64}",
                   "inds": inds, "data": data}, f, indent=1)


if __name__ == "__main__":
    main()
