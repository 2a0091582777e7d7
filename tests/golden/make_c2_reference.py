"""Oracle reference for the full-size Poiseuille workload (192x129x192), 20 steps.

The oracle needs ~2 minutes of CPU for this, too long for a GPU test, so its
result is committed as a compact fixture: per-component L2 norms of u and q
after 20 steps, and every Chebyshev coefficient of a fixed set of Fourier
modes. Generated with the oracle (oracle/libcpuref.so) in this repository.

Run: python tests/golden/make_c2_reference.py
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
from test_gpu_parity import C2_WORKLOAD  # noqa: E402

MODES = [(0, 0), (1, 0), (-1, 0), (0, 1), (1, 1), (-1, 1), (2, 3), (-5, 7), (30, 10), (-63, 63), (63, 0), (17, 41)]
STEPS = 20


def main():
    wl = dict(C2_WORKLOAD, label="plane Poiseuille Re_bulk=2800, 192x129x192, rotational NL, SBDF3 dt=0.005",
              norm=0.2, decay=0.5, seed=2024)
    Nx, Ny, Nz = wl["grid"]
    g = R.Grid(Nx, Ny, Nz, wl["Lx"], wl["Lz"])
    u0 = random_field(Nx, Ny, Nz, wl["Lx"], wl["Lz"], norm=wl["norm"], decay=wl["decay"], seed=wl["seed"])
    d = R.OracleDNS(g, R.make_flags(**wl["flags"]), u0, g.zeros(1))
    d.advance(STEPS)
    u, q = d.state()
    norms = [float(np.sqrt(R.l2norm2(g, u[i:i + 1]))) for i in range(3)]
    qn = float(np.sqrt(R.l2norm2(g, q)))
    prof = np.stack([u[:, :, kx % Nx, kz] for kx, kz in MODES], axis=0)  # (nmodes, 3, Ny)
    prof.astype(np.complex128).tofile(os.path.join(HERE, "c2_20steps_modes.c128"))
    meta = dict(workload=wl["label"], steps=STEPS, modes=MODES, shape=list(prof.shape), l2norm_u=norms,
                l2norm_q=qn, time=d.time, generator="oracle/cpu_ref.cpp via tests/cfref.py",
                seed=wl["seed"], norm=wl["norm"], decay=wl["decay"])
    json.dump(meta, open(os.path.join(HERE, "c2_20steps.json"), "w"), indent=1)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
