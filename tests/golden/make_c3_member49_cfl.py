"""Writes c3_member49_cfl.json: member 49 of the C3 ensemble (64x65x64 plane
Couette at Re 400, skew-symmetric, SBDF3 with dt = 0.02, initial norm 0.27)
integrated by the oracle, its CFL number (DNS::CFL) every 25 steps, and the
first multiple of 25 steps after which its velocity is no longer finite. At
dt = 0.02 this member's CFL number climbs past 0.7 and SBDF3 diverges: the
fixture pins that the divergence belongs to the scheme at this dt, not to a
GPU kernel. About 70 s on one core.

    python tests/golden/make_c3_member49_cfl.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))
import bench  # noqa: E402
import cfref as R  # noqa: E402

MEMBER, CHUNK, MAX_STEPS = 49, 25, 400


def main():
    m = bench.c3_members()[MEMBER]
    Nx, Ny, Nz = m["grid"]
    g = R.Grid(Nx, Ny, Nz, bench.LX_C, bench.LZ_C)
    d = R.OracleDNS(g, R.make_flags(**m["flags"]), bench.member_field(m), g.zeros(1))
    steps, cfl, bad = [], [], None
    for s in range(CHUNK, MAX_STEPS + 1, CHUNK):
        d.advance(CHUNK)
        u, _ = d.state()
        if not np.all(np.isfinite(u)):
            bad = s
            break
        steps.append(s)
        cfl.append(d.cfl())
    out = dict(member=MEMBER, grid=list(m["grid"]), dt=m["flags"]["dt"], chunk=CHUNK, steps=steps, cfl=cfl,
               nonfinite_after_step=bad)
    with open(os.path.join(HERE, "c3_member49_cfl.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
