"""Oracle references for the two largest serving grid classes, 20 steps each.

  96x97x96    plane Poiseuille, Re_bulk = 3000, rotational, dt = 0.01 * 32/96
  128x129x128 plane Couette, Re = 400, skew-symmetric, bulk velocity, dt = 0.02 * 32/128

(the serving workload's flows and time steps: bench.py serve_requests), each
from a seeded divergence-free perturbation of norm 0.3, the top of the
serving range. The oracle needs about half a minute per case, so the results
are committed as compact fixtures, as for the 192x129x192 channel: the
per-component L2 norms of u and q after 20 steps, and every Chebyshev
coefficient of a fixed set of Fourier modes. Generated with the oracle
(oracle/libcpuref.so) of this repository.

Run: python tests/golden/make_c5_reference.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cfref as R  # noqa: E402
from channelflow_amd.fields import random_field  # noqa: E402

LX, LZ = 2 * np.pi / 1.14, 2 * np.pi / 2.5
STEPS = 20
CASES = {
    "c5_96_poiseuille": dict(grid=(96, 97, 96), norm=0.3, decay=0.3, seed=2024, member=96,
                             flags=dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 / 3000,
                                        nu=1.0 / 3000, timestepping="SBDF3", initstepping="SMRK2",
                                        nonlinearity="Rotational", dealiasing="DealiasXZ", taucorrection=True,
                                        dt=0.01 * 32 / 96)),
    "c5_128_couette": dict(grid=(128, 129, 128), norm=0.3, decay=0.3, seed=2024, member=128,
                           flags=dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0,
                                      constraint="BulkVelocity", Ubulk=0.0, nu=1.0 / 400, timestepping="SBDF3",
                                      initstepping="SMRK2", nonlinearity="SkewSymmetric", dealiasing="DealiasXZ",
                                      taucorrection=True, dt=0.02 * 32 / 128)),
}
MODES = [(0, 0), (1, 0), (-1, 0), (0, 1), (1, 1), (-1, 1), (2, 3), (-5, 7), (20, 10), (-30, 30), (30, 0), (11, 25)]


def initial(case):
    Nx, Ny, Nz = case["grid"]
    return random_field(Nx, Ny, Nz, LX, LZ, norm=case["norm"], decay=case["decay"], seed=case["seed"],
                        member=case["member"])


def main(names):
    for name in names:
        case = CASES[name]
        Nx, Ny, Nz = case["grid"]
        g = R.Grid(Nx, Ny, Nz, LX, LZ)
        d = R.OracleDNS(g, R.make_flags(**case["flags"]), initial(case), g.zeros(1))
        d.advance(STEPS)
        u, q = d.state()
        prof = np.stack([u[:, :, kx % Nx, kz] for kx, kz in MODES], axis=0)  # (nmodes, 3, Ny)
        prof.astype(np.complex128).tofile(os.path.join(HERE, name + "_modes.c128"))
        meta = dict(case=name, grid=list(case["grid"]), flags=case["flags"], steps=STEPS, modes=MODES,
                    shape=list(prof.shape), l2norm_u=[float(np.sqrt(R.l2norm2(g, u[i:i + 1]))) for i in range(3)],
                    l2norm_q=float(np.sqrt(R.l2norm2(g, q))), time=d.time,
                    generator="oracle/cpu_ref.cpp via tests/cfref.py", seed=case["seed"], member=case["member"],
                    norm=case["norm"], decay=case["decay"])
        json.dump(meta, open(os.path.join(HERE, name + ".json"), "w"), indent=1)
        print(name, meta["l2norm_u"], flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
