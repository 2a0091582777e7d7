"""C ABI checks that need no GPU: the library loads, exports every entry
point declared in include/cfgpu.h, and reports errors as status codes."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cfref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cfgpu.h")
LIB = os.path.join(ROOT, "channelflow_amd", "libcfgpu.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cfg_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("cfg_dns_create", "cfg_dns_advance", "cfg_dns_get_state", "cfg_dns_set_state", "cfg_nonlinear",
                 "cfg_solve", "cfg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (python -m channelflow_amd.build)"
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    for n in declared():
        assert re.search(r"\bT " + n + r"\b", out), n


def test_python_mirror_matches_header():
    import channelflow_amd as cfa
    assert sorted(cfa.EXPORTED) == declared()


def test_errors_are_status_codes():
    import channelflow_amd as cfa
    lib = cfa.library()
    # invalid flags are rejected before any device work
    f = cfa.DNSFlags(dealiasing="DealiasY").c()
    g = cfa.FlowField(8, 9, 8, 3, 1.0, 1.0).grid()
    u = np.zeros((3, 9, 8, 5), dtype=np.complex128)
    h = C.c_void_p()
    rc = lib.cfg_dns_create(C.byref(g), C.byref(f), u.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                            None, None, None, C.byref(h))
    assert rc == -1 and b"DealiasY" in lib.cfg_last_error()
    f = cfa.DNSFlags().c()
    g = cfa.FlowField(8, 10, 8, 3, 1.0, 1.0).grid()  # even Ny: the tau solver needs odd Ny
    rc = lib.cfg_dns_create(C.byref(g), C.byref(f), u.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                            None, None, None, C.byref(h))
    assert rc == -1 and b"odd Ny" in lib.cfg_last_error()


def test_host_build_of_per_mode_math_matches_oracle(tmp_path):
    """cfg_math.hpp (the code the tau kernels run) compiled for the host equals the oracle bit for bit."""
    so = tmp_path / "libhostmath.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(so),
                    os.path.join(ROOT, "tests", "hostmath", "cfg_hostmath.cpp")], check=True)
    H = C.CDLL(str(so))
    P = C.POINTER(C.c_double)
    H.hm_tausolve.argtypes = [C.c_int, C.c_int] + [C.c_double] * 6 + [C.c_int, C.c_int, P, P]
    rng = np.random.default_rng(1)
    Ny, Lx, Lz, nu, lt = 33, 2 * np.pi / 1.14, 2 * np.pi / 2.5, 1 / 400, 11 / 6 / 0.02
    for kx, kz in [(1, 0), (0, 1), (3, 5), (-4, 2), (0, 0), (10, 9)]:
        Rv = rng.standard_normal((3, Ny)) + 1j * rng.standard_normal((3, Ny))
        lam = lt + 4 * np.pi ** 2 * nu * ((kx / Lx) ** 2 + (kz / Lz) ** 2)
        ref = R.tausolve(kx, kz, Lx, Lz, -1, 1, lam, nu, Ny, True, Rv[0], Rv[1], Rv[2])
        Rin = np.ascontiguousarray(Rv.reshape(-1)).view(np.float64)
        out = np.zeros(8 * Ny)
        H.hm_tausolve(kx, kz, Lx, Lz, -1.0, 1.0, lt, nu, Ny, 1, Rin.ctypes.data_as(P), out.ctypes.data_as(P))
        got = out.view(np.complex128).reshape(4, Ny)
        assert np.array_equal(got, ref), (kx, kz, np.abs(got - ref).max())
