"""Convert channelflow's Orr-Sommerfeld eigenfunction fixtures to JSON.

Source (read in place, never copied): ``tests/data/os_ueig10_65.asc``,
``tests/data/os_peig10_65.asc`` and ``tests/data/os_omega10_65.cmplx`` of the
reference checkout, the inputs of ``tests/dnsOrrsommTest.cpp:163-176``.

The ``.asc`` files are BasisFunc text files (``channelflow/basisfunc.cpp:43-73``):
a header ``% Nd Ny kx kz Lx Lz a b state`` and then, for each of the Ny
Chebyshev points y_n = cos(pi n/(Ny-1)) (state ``P``: physical values), the
real and imaginary part of every component. The ``.cmplx`` file holds the
eigenvalue omega (real, imaginary); the mode grows as exp(-i omega t).

Output: ``os_eig10_65.json`` with the headers, the values as
``[component][n] = [re, im]`` and omega, printed with 17 significant digits so
that every double survives the round trip.

Run:  python tests/golden/make_os_fixture.py [/root/reference/tests/data]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def read_basisfunc(path):
    with open(path) as fh:
        head = fh.readline().split()
        if head[0] != "%":
            raise ValueError(f"bad BasisFunc header in {path}")
        Nd, Ny, kx, kz = (int(v) for v in head[1:5])
        Lx, Lz, a, b = (float(v) for v in head[5:9])
        state = head[9]
        vals = [float(v) for v in fh.read().split()]
    if len(vals) != 2 * Nd * Ny:
        raise ValueError(f"{path}: expected {2 * Nd * Ny} numbers, found {len(vals)}")
    comps = [[[vals[2 * (ny * Nd + c)], vals[2 * (ny * Nd + c) + 1]] for ny in range(Ny)] for c in range(Nd)]
    return dict(Nd=Nd, Ny=Ny, kx=kx, kz=kz, Lx=Lx, Lz=Lz, a=a, b=b, state=state, values=comps)


def convert(datadir):
    out = dict(source="tests/data/os_{u,p}eig10_65.asc, os_omega10_65.cmplx (channelflow 2.0.2)",
               cite="tests/dnsOrrsommTest.cpp:39-81,163-179",
               u=read_basisfunc(os.path.join(datadir, "os_ueig10_65.asc")),
               p=read_basisfunc(os.path.join(datadir, "os_peig10_65.asc")))
    with open(os.path.join(datadir, "os_omega10_65.cmplx")) as fh:
        re_, im_ = (float(v) for v in fh.read().split())
    out["omega"] = [re_, im_]
    dst = os.path.join(HERE, "os_eig10_65.json")
    with open(dst, "w") as fh:
        fh.write(json.dumps(out, indent=None, separators=(",", ":"),
                            default=float).replace("], [", "],["))
    # re-read and check every value round-trips exactly
    back = json.load(open(dst))
    for k in ("u", "p"):
        assert back[k]["values"] == out[k]["values"]
    print("wrote", dst)


if __name__ == "__main__":
    convert(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data")
