"""Convert channelflow's golden trajectory fixtures to raw little-endian FP64.

Source (read in place, never copied): ``tests/data/uinit.nc`` and
``tests/data/ufinal.nc`` of the reference checkout, the start and end field of
``tests/timeIntegrationTest.cpp:42,54,118`` (48x35x48 grid, 440 SBDF3 steps,
pass bar ``L2Dist <= 1e-13``).

Each NetCDF4 file stores the *de-padded physical* velocity: datasets
``Velocity_X/Y/Z`` of shape (Z, Y, X) = (32, 35, 32) with X fastest
(``flowfield.cpp:3797-3806``), on the grid ``Nx_io = 2*(Nx/3)``,
``Nz_io = 2*(Nz/3)`` (``flowfield.cpp:520-521``), y at the Chebyshev points
``cos(pi*ny/(Ny-1))`` from +1 down to -1.

The HDF5 reader used here is ``h5dump`` (conda's HDF5 tools, present only in
the build container).  Output: ``<name>.f64`` holding the three components
back to back, each (Z, Y, X) C-order, plus ``manifest.json`` with the grid
attributes (printed at 17 significant digits) and a sha256 per file.

Run:  python tests/golden/make_fixtures.py [/root/reference/tests/data]
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))


def _attr(path, name):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-a", "/" + name, path],
                         check=True, capture_output=True, text=True).stdout
    m = re.search(r"\(0\):\s*([-+0-9.eE]+)", out)
    if not m:
        raise RuntimeError(f"attribute {name} not found in {path}")
    return float(m.group(1))


def _dataset(path, name, shape):
    with tempfile.TemporaryDirectory() as td:
        raw = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-b", "LE", "-d", "/" + name, "-o", raw, path],
                       check=True, capture_output=True)
        arr = np.fromfile(raw, dtype="<f8")
    return arr.reshape(shape)


def convert(datadir):
    manifest = {}
    for base in ("uinit", "ufinal"):
        src = os.path.join(datadir, base + ".nc")
        attrs = {k: _attr(src, k) for k in ("Nx", "Ny", "Nz", "Lx", "Lz", "a", "b")}
        nx_io = 2 * (int(attrs["Nx"]) // 3)
        nz_io = 2 * (int(attrs["Nz"]) // 3)
        ny = int(attrs["Ny"])
        comps = [_dataset(src, "Velocity_" + c, (nz_io, ny, nx_io)) for c in "XYZ"]
        data = np.stack(comps).astype("<f8")
        out = os.path.join(HERE, base + ".f64")
        data.tofile(out)
        with open(out, "rb") as fh:
            digest = hashlib.sha256(fh.read()).hexdigest()
        manifest[base] = {
            "file": base + ".f64",
            "layout": "component, z, y, x (x fastest); de-padded physical grid",
            "shape": [3, nz_io, ny, nx_io],
            "Nx": int(attrs["Nx"]), "Ny": ny, "Nz": int(attrs["Nz"]),
            "Lx": attrs["Lx"], "Lz": attrs["Lz"], "a": attrs["a"], "b": attrs["b"],
            "sha256": digest,
        }
    manifest["timeIntegrationTest"] = {
        "cite": "tests/timeIntegrationTest.cpp:71-98,104-137",
        "baseflow": "SuctionBase", "timestepping": "SBDF3", "initstepping": "SMRK2",
        "nonlinearity": "Rotational", "dealiasing": "DealiasXZ", "taucorrection": True,
        "constraint": "PressureGradient", "dPdx": 0.0, "dt": 1.0 / 40, "nu": 1.0 / 400,
        "Vsuck": 1.0 / 400, "steps": 440, "tolerance_L2Dist": 1e-13,
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    return manifest


if __name__ == "__main__":
    d = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data"
    print(json.dumps(convert(d), indent=1))
