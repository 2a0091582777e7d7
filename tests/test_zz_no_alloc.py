"""cfg_nonlinear_noalloc (include/cfgpu.h), the entry point that allocates no
device memory, against cfg_nonlinear, the allocating entry point it varies:
the same NSE::nonlinear evaluation (nse.cpp:383-391), so the live modes must
agree bit for bit. Calls in a row on one device, the second on a smaller
grid than the first (the shared scratch array then holds the first call's
data), the last on 64x65x64, the largest C3 grid, whose z pass stages its
inputs in LDS; and a grid beyond the scratch array's size rejected with a
status code before any device work."""
import ctypes as C

import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu

LX, LZ = 2 * np.pi / 1.14, 2 * np.pi / 2.5
P = C.POINTER(C.c_double)


def live_index(Nx, Nz, dealias):
    kxd = Nx // 3 - 1 if dealias else Nx // 2 - 1
    kzd = Nz // 3 - 1 if dealias else Nz // 2 - 1
    nkx = 2 * kxd + 1
    mx = np.array([(i if i <= kxd else i - nkx) % Nx for i in range(nkx)])
    return mx, kzd + 1


def to_compact(u, mx, nkz):
    """(3, Ny, Nx, Nz/2+1) reference layout -> (3, Ny, nkz, nkx) compact layout."""
    return np.ascontiguousarray(u[:, :, mx, :nkz].transpose(0, 1, 3, 2))


def hip():
    """The HIP runtime the library itself runs on (the test's device buffers
    come from it; the entry point under test allocates nothing)."""
    cfa.library()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def call_noalloc(grid, flags, uc):
    h = hip()
    n = uc.nbytes
    ud, fd = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(ud), n) == 0 and h.hipMalloc(C.byref(fd), n) == 0
    try:
        assert h.hipMemcpy(ud, uc.ctypes.data_as(C.c_void_p), n, H2D) == 0
        assert h.hipMemset(fd, 0xFF, n) == 0  # NaN pattern: every live value must be written
        assert h.hipDeviceSynchronize() == 0
        lib = cfa.library()
        rc = lib.cfg_nonlinear_noalloc(C.byref(grid), C.byref(flags), C.cast(ud, P), C.cast(fd, P), None)
        assert rc == 0, lib.cfg_last_error().decode()
        out = np.empty_like(uc)
        assert h.hipMemcpy(out.ctypes.data_as(C.c_void_p), fd, n, D2H) == 0
    finally:
        h.hipFree(ud)
        h.hipFree(fd)
    return out


@pytest.mark.parametrize("calls", [[((48, 49, 48), "SkewSymmetric", "DealiasXZ"),
                                    ((16, 17, 16), "Rotational", "DealiasXZ"),
                                    ((24, 25, 18), "Convection", "NoDealiasing"),
                                    ((64, 65, 64), "SkewSymmetric", "DealiasXZ")]])
def test_noalloc_matches_nonlinear(gpu, calls):
    for k, (shape, method, dealias) in enumerate(calls):
        Nx, Ny, Nz = shape
        print(f"call {k}: {Nx}x{Ny}x{Nz} {method} {dealias}", flush=True)
        u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.2, decay=0.4, seed=40 + k, dealias=(dealias == "DealiasXZ"))
        fl = cfa.DNSFlags(nonlinearity=method, dealiasing=dealias, baseflow="LaminarBase", ulowerwall=-1.0,
                          uupperwall=1.0, constraint="BulkVelocity", nu=1 / 400)
        ff = cfa.FlowField(Nx, Ny, Nz, 3, LX, LZ, data=u)
        f_ref = cfa.nonlinear(ff, fl)
        mx, nkz = live_index(Nx, Nz, dealias == "DealiasXZ")
        f_na = call_noalloc(ff.grid(), fl.c(), to_compact(u, mx, nkz))
        ref = to_compact(f_ref, mx, nkz)
        assert np.all(np.isfinite(f_na))
        assert np.array_equal(f_na, ref), np.max(np.abs(f_na - ref))


def test_noalloc_rejects_oversized_grid(gpu):
    g = cfa.FlowField(96, 97, 96, 3, LX, LZ).grid()
    fl = cfa.DNSFlags(nonlinearity="SkewSymmetric").c()
    dummy = np.zeros(2, dtype=np.float64)
    lib = cfa.library()
    rc = lib.cfg_nonlinear_noalloc(C.byref(g), C.byref(fl), dummy.ctypes.data_as(P), dummy.ctypes.data_as(P), None)
    assert rc == -1 and b"scratch" in lib.cfg_last_error()
