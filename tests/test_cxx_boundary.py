"""The compiled side of the boundary: tests/cxx/dns_driver.bin, a C++ program
built with g++ against include/cfgpu.hpp (channelflow's DNS signatures,
channelflow/dns.h:30-97) and libcfgpu.so.

CPU: the facade header compiles on its own, and the driver loads the library
and resolves its symbols. GPU: the driver runs C1's physics (Couette Re 400,
skew-symmetric, bulk velocity, SBDF3 from SMRK2, dt 0.02) at 32x33x32 for 20
steps, as advance() calls of 5 with CFL and time between them, and matches the
oracle per velocity component at 1e-10; a DNSBatch member gives the same
bits as the DNS."""
import os
import subprocess

import numpy as np
import pytest

import cfref as R
from channelflow_amd.fields import random_field

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(HERE, "cxx", "dns_driver.bin")
LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5
C1_FLAGS = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
                nu=1 / 400, dt=0.02, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric",
                dealiasing="DealiasXZ")


def _driver():
    if not os.path.exists(DRIVER):
        from channelflow_amd import build
        build.build_driver()
    return DRIVER


def test_facade_header_compiles_alone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "cfgpu.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                    os.path.join(ROOT, "include"), str(src)], check=True)


def test_driver_loads_library():
    r = subprocess.run([_driver(), "probe"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "cfgpu loaded" in r.stdout


def _run(mode, u0, grid, nsteps, tmp_path):
    Nx, Ny, Nz = grid
    fin, fout = tmp_path / f"{mode}_in.bin", tmp_path / f"{mode}_out.bin"
    np.ascontiguousarray(u0, dtype=np.complex128).tofile(fin)
    print("driver", mode, grid, nsteps, flush=True)
    r = subprocess.run([_driver(), mode, str(Nx), str(Ny), str(Nz), repr(LX_C), repr(LZ_C), str(nsteps), str(fin),
                        str(fout)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return np.fromfile(fout, dtype=np.complex128).reshape(u0.shape), r.stdout


@pytest.mark.gpu
def test_cxx_driver_c1_20_steps_vs_oracle(gpu, tmp_path):
    grid = (32, 33, 32)
    u0 = random_field(*grid, LX_C, LZ_C, norm=0.1, decay=0.3, seed=1)
    ug, out = _run("dns", u0, grid, 20, tmp_path)
    assert "t = 0.400" in out
    g = R.Grid(*grid, LX_C, LZ_C)
    ref = R.OracleDNS(g, R.make_flags(**C1_FLAGS), u0, g.zeros(1))
    ref.advance(20)
    ur, _ = ref.state()
    for i in range(3):
        d = np.sqrt(R.l2dist2(g, ug[i:i + 1], ur[i:i + 1]))
        n = np.sqrt(R.l2norm2(g, ur[i:i + 1]))
        print("component", i, "rel L2", d / n, flush=True)
        assert d / n <= 1e-10
    ub, _ = _run("batch", u0, grid, 20, tmp_path)
    assert np.array_equal(ub, ug)
