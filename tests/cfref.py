"""ctypes binding of the oracle (oracle/libcpuref.so) for the test suite.

The oracle is test infrastructure: this module is imported by tests/ only (and
by bench.py's cpu_baseline leg through its own loader).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# CFREF_LIB points the tests at another build of the oracle (one with a planted
# mistake, to show that the pins catch it)
LIB_PATH = os.environ.get("CFREF_LIB") or os.path.join(ROOT, "oracle", "libcpuref.so")

# enums numbered as channelflow/dnsflags.h:26-38
BASEFLOW = dict(ZeroBase=0, LinearBase=1, ParabolicBase=2, LaminarBase=3, SuctionBase=4, ArbitraryBase=5)
CONSTRAINT = dict(PressureGradient=0, BulkVelocity=1)
STEPPING = dict(CNFE1=0, CNAB2=1, CNRK2=2, SMRK2=3, SBDF1=4, SBDF2=5, SBDF3=6, SBDF4=7)
NONLINEARITY = dict(Rotational=0, Convection=1, Divergence=2, SkewSymmetric=3, Alternating=4, Alternating_=5)
DEALIASING = dict(NoDealiasing=0, DealiasXZ=1, DealiasY=2, DealiasXYZ=3)


class Flags(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("baseflow", "constraint", "timestepping", "initstepping",
                                       "nonlinearity", "dealiasing", "taucorrection", "reserved")] + \
               [(n, C.c_double) for n in ("nu", "Vsuck", "rotation", "dPdx", "dPdz", "Ubulk", "Wbulk",
                                          "ulowerwall", "uupperwall", "wlowerwall", "wupperwall", "t0", "dt")]


def make_flags(**kw):
    """Flags with channelflow's DNSFlags defaults (dnsflags.h:84-92)."""
    d = dict(baseflow="LaminarBase", constraint="PressureGradient", timestepping="SBDF3",
             initstepping="SMRK2", nonlinearity="Rotational", dealiasing="DealiasXZ",
             taucorrection=True, nu=0.0025, Vsuck=0.0, rotation=0.0, dPdx=0.0, dPdz=0.0, Ubulk=0.0,
             Wbulk=0.0, ulowerwall=0.0, uupperwall=0.0, wlowerwall=0.0, wupperwall=0.0, t0=0.0,
             dt=0.03125)
    d.update(kw)
    f = Flags()
    f.baseflow = BASEFLOW[d["baseflow"]] if isinstance(d["baseflow"], str) else d["baseflow"]
    f.constraint = CONSTRAINT[d["constraint"]] if isinstance(d["constraint"], str) else d["constraint"]
    f.timestepping = STEPPING[d["timestepping"]] if isinstance(d["timestepping"], str) else d["timestepping"]
    f.initstepping = STEPPING[d["initstepping"]] if isinstance(d["initstepping"], str) else d["initstepping"]
    f.nonlinearity = NONLINEARITY[d["nonlinearity"]] if isinstance(d["nonlinearity"], str) else d["nonlinearity"]
    f.dealiasing = DEALIASING[d["dealiasing"]] if isinstance(d["dealiasing"], str) else d["dealiasing"]
    f.taucorrection = int(bool(d["taucorrection"]))
    for n in ("nu", "Vsuck", "rotation", "dPdx", "dPdz", "Ubulk", "Wbulk", "ulowerwall", "uupperwall",
              "wlowerwall", "wupperwall", "t0", "dt"):
        setattr(f, n, float(d[n]))
    return f


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(C.c_double)
        i, d, v = C.c_int, C.c_double, C.c_void_p
        F = C.POINTER(Flags)
        sig = {
            "cpuref_last_error": ([], C.c_char_p),
            "cpuref_dns_create": ([i, i, i, d, d, d, d, F, P, P, i, P, P], v),
            "cpuref_dns_advance": ([v, i], i),
            "cpuref_dns_reset_dt": ([v, d], i),
            "cpuref_dns_get": ([v, P, P], None),
            "cpuref_dns_set": ([v, P, P], None),
            "cpuref_dns_time": ([v], d),
            "cpuref_dns_dpdx": ([v], d),
            "cpuref_dns_ubulk": ([v], d),
            "cpuref_dns_cfl": ([v], d),
            "cpuref_dns_base": ([v, P, P], None),
            "cpuref_dns_destroy": ([v], None),
            "cpuref_nonlinear": ([i, i, i, d, d, d, d, F, P, P], i),
            "cpuref_linear": ([i, i, i, d, d, d, d, F, P, P, P], i),
            "cpuref_solve": ([i, i, i, d, d, d, d, F, d, P, P, P, P, P], i),
            "cpuref_tausolve": ([i, i, d, d, d, d, d, d, i, i, P, P], i),
            "cpuref_tausolve_bulk": ([d, d, d, d, d, d, i, i, P, d, d, P, P], i),
            "cpuref_helmholtz": ([i, d, d, d, d, P, d, d, P], i),
            "cpuref_helmholtz_mean": ([i, d, d, d, d, P, d, d, d, P, P], i),
            "cpuref_banded_solve": ([i, P, P, P, P, P], i),
            "cpuref_cheb_diff": ([i, d, d, P, P], None),
            "cpuref_cheb_spectral": ([i, P], None),
            "cpuref_cheb_physical": ([i, P], None),
            "cpuref_cheb_mean": ([i, P], d),
            "cpuref_cheb_l2norm2": ([i, d, d, P], d),
            "cpuref_field_make_spectral": ([i, i, i, i, d, d, d, d, P], None),
            "cpuref_field_make_physical": ([i, i, i, i, d, d, d, d, P], None),
            "cpuref_l2norm2": ([i, i, i, i, d, d, d, d, P, i], d),
            "cpuref_l2dist2": ([i, i, i, i, d, d, d, d, P, P, i], d),
            "cpuref_field_from_io": ([i, i, i, i, d, d, d, d, P, P], i),
            "cpuref_pressure": ([i, i, i, d, d, d, d, F, P, P], i),
            "cpuref_laminar_profile": ([d, i, d, d, d, d, d, d, d, i, P], i),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().cpuref_last_error().decode())


class Grid:
    """Grid geometry; spectral arrays are complex128 with shape (Nd, Ny, Nx, Nz//2+1)."""

    def __init__(self, Nx, Ny, Nz, Lx, Lz, a=-1.0, b=1.0):
        self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b = Nx, Ny, Nz, float(Lx), float(Lz), float(a), float(b)
        self.Mz = Nz // 2 + 1

    def args(self):
        return (self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b)

    def zeros(self, nd=3):
        return np.zeros((nd, self.Ny, self.Nx, self.Mz), dtype=np.complex128)


def _as_f64(c):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    return c.view(np.float64)


class OracleDNS:
    """The oracle's DNS (dns.cpp) on spectral arrays in channelflow's layout."""

    def __init__(self, grid, flags, u, q=None, padded=True, base=None):
        self.g = grid
        uf = _as_f64(u).copy()
        qf = _as_f64(q).copy() if q is not None else None
        U = W = None
        if base is not None:
            U = np.ascontiguousarray(base[0], dtype=np.float64)
            W = np.ascontiguousarray(base[1], dtype=np.float64)
        self.h = lib().cpuref_dns_create(*grid.args(), C.byref(flags), ptr(uf), ptr(qf), int(padded),
                                         ptr(U), ptr(W))
        if not self.h:
            raise RuntimeError("oracle: " + lib().cpuref_last_error().decode())

    def advance(self, n):
        check(lib().cpuref_dns_advance(self.h, n))

    def reset_dt(self, dt):
        check(lib().cpuref_dns_reset_dt(self.h, dt))

    def set_state(self, u, q=None):
        uf = _as_f64(u).copy()
        qf = _as_f64(q).copy() if q is not None else None
        lib().cpuref_dns_set(self.h, ptr(uf), ptr(qf))

    def state(self):
        u = self.g.zeros(3)
        q = self.g.zeros(1)
        lib().cpuref_dns_get(self.h, ptr(u.view(np.float64)), ptr(q.view(np.float64)))
        return u, q

    def base(self):
        U = np.zeros(self.g.Ny)
        W = np.zeros(self.g.Ny)
        lib().cpuref_dns_base(self.h, ptr(U), ptr(W))
        return U, W

    time = property(lambda self: lib().cpuref_dns_time(self.h))
    dPdx = property(lambda self: lib().cpuref_dns_dpdx(self.h))
    Ubulk = property(lambda self: lib().cpuref_dns_ubulk(self.h))

    def cfl(self):
        return lib().cpuref_dns_cfl(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().cpuref_dns_destroy(self.h)
            self.h = None


def pressure(grid, flags, u):
    """PressureSolver(u, Ubase, Wbase, nu, Vsuck, nonlinearity).solve(p, u) with the
    base flow of a DNS built from the same flags (simulateflow.cpp:100-105)."""
    uf = _as_f64(u)
    p = np.zeros((1, grid.Ny, grid.Nx, grid.Nz // 2 + 1), dtype=np.complex128)
    check(lib().cpuref_pressure(*grid.args(), C.byref(flags), ptr(uf), ptr(p.view(np.float64))))
    return p


def nonlinear(grid, flags, u):
    uf = _as_f64(u).copy()
    f = grid.zeros(3)
    check(lib().cpuref_nonlinear(*grid.args(), C.byref(flags), ptr(uf), ptr(f.view(np.float64))))
    return f, uf.view(np.complex128).reshape(u.shape)


def linear(grid, flags, u, q):
    uf = _as_f64(u).copy()
    qf = _as_f64(q).copy()
    out = grid.zeros(3)
    check(lib().cpuref_linear(*grid.args(), C.byref(flags), ptr(uf), ptr(qf), ptr(out.view(np.float64))))
    return out


def solve(grid, flags, lambda_t, u0, rhs, u_init=None, q_init=None):
    u0f = _as_f64(u0).copy()
    rf = _as_f64(rhs).copy()
    uo = grid.zeros(3) if u_init is None else np.ascontiguousarray(u_init, dtype=np.complex128).copy()
    qo = grid.zeros(1) if q_init is None else np.ascontiguousarray(q_init, dtype=np.complex128).copy()
    dp = np.zeros(1)
    check(lib().cpuref_solve(*grid.args(), C.byref(flags), float(lambda_t), ptr(u0f), ptr(rf),
                             ptr(uo.view(np.float64)), ptr(qo.view(np.float64)), ptr(dp)))
    return uo, qo, float(dp[0])


def tausolve(kx, kz, Lx, Lz, a, b, lam, nu, Ny, taucorr, Rx, Ry, Rz):
    R = np.ascontiguousarray(np.concatenate([Rx, Ry, Rz]).astype(np.complex128)).view(np.float64)
    out = np.zeros(4 * Ny, dtype=np.complex128)
    check(lib().cpuref_tausolve(kx, kz, Lx, Lz, a, b, lam, nu, Ny, int(taucorr), ptr(R), ptr(out.view(np.float64))))
    return out.reshape(4, Ny)


def tausolve_bulk(Lx, Lz, a, b, lam, nu, Ny, taucorr, Rx, Ry, Rz, umean, wmean):
    """TauSolver::solve with dPdx, dPdz unknown and mean(u), mean(w) imposed (kx = kz = 0)."""
    R = np.ascontiguousarray(np.concatenate([Rx, Ry, Rz]).astype(np.complex128)).view(np.float64)
    out = np.zeros(4 * Ny, dtype=np.complex128)
    dP = np.zeros(2)
    check(lib().cpuref_tausolve_bulk(Lx, Lz, a, b, lam, nu, Ny, int(taucorr), ptr(R), umean, wmean,
                                     ptr(out.view(np.float64)), ptr(dP)))
    return out.reshape(4, Ny), dP[0], dP[1]


def field_from_io(grid, io):
    out = grid.zeros(3)
    io = np.ascontiguousarray(io, dtype=np.float64)
    check(lib().cpuref_field_from_io(grid.Nx, grid.Ny, grid.Nz, io.shape[0], grid.Lx, grid.Lz, grid.a, grid.b,
                                     ptr(io), ptr(out.view(np.float64))))
    return out


def l2norm2(grid, u, padded=True):
    uf = _as_f64(u)
    return lib().cpuref_l2norm2(grid.Nx, grid.Ny, grid.Nz, u.shape[0], grid.Lx, grid.Lz, grid.a, grid.b,
                                ptr(uf), int(padded))


def l2dist2(grid, u, v, padded=True):
    return lib().cpuref_l2dist2(grid.Nx, grid.Ny, grid.Nz, u.shape[0], grid.Lx, grid.Lz, grid.a, grid.b,
                                ptr(_as_f64(u)), ptr(_as_f64(v)), int(padded))


def field_physical(grid, u):
    """Spectral (Nd,Ny,Nx,Mz) complex -> physical (Nd,Ny,Nx,Nz) real."""
    nd = u.shape[0]
    buf = _as_f64(u).copy()
    lib().cpuref_field_make_physical(grid.Nx, grid.Ny, grid.Nz, nd, grid.Lx, grid.Lz, grid.a, grid.b, ptr(buf))
    return buf.reshape(nd, grid.Ny, grid.Nx, 2 * grid.Mz)[..., :grid.Nz].copy()


def field_spectral(grid, phys):
    nd = phys.shape[0]
    buf = np.zeros((nd, grid.Ny, grid.Nx, 2 * grid.Mz))
    buf[..., :grid.Nz] = phys
    lib().cpuref_field_make_spectral(grid.Nx, grid.Ny, grid.Nz, nd, grid.Lx, grid.Lz, grid.a, grid.b, ptr(buf))
    return buf.view(np.complex128).reshape(nd, grid.Ny, grid.Nx, grid.Mz).copy()
