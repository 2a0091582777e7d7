"""channelflow_amd -- MI355X-native DNS time step for channelflow.

Python host mirror of channelflow's DNSFlags / FlowField / DNS interface
(channelflow/dnsflags.h:81-145, flowfield.h:52-345, dns.h:30-97) over the C
ABI of libcfgpu.so (include/cfgpu.h). All time stepping runs in the HIP
kernels of the library; this module only marshals flags and spectral arrays.
There is no CPU fallback: on a machine without the library or without a GPU
every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfgpu.so")

# enums numbered as channelflow/dnsflags.h:26-38
BASEFLOW = dict(ZeroBase=0, LinearBase=1, ParabolicBase=2, LaminarBase=3, SuctionBase=4, ArbitraryBase=5)
CONSTRAINT = dict(PressureGradient=0, BulkVelocity=1)
STEPPING = dict(CNFE1=0, CNAB2=1, CNRK2=2, SMRK2=3, SBDF1=4, SBDF2=5, SBDF3=6, SBDF4=7)
NONLINEARITY = dict(Rotational=0, Convection=1, Divergence=2, SkewSymmetric=3, Alternating=4, Alternating_=5)
DEALIASING = dict(NoDealiasing=0, DealiasXZ=1, DealiasY=2, DealiasXYZ=3)
_ENUMS = dict(baseflow=BASEFLOW, constraint=CONSTRAINT, timestepping=STEPPING, initstepping=STEPPING,
              nonlinearity=NONLINEARITY, dealiasing=DEALIASING)

EXPORTED = ("cfg_last_error", "cfg_device_count", "cfg_dns_create", "cfg_dns_destroy", "cfg_dns_advance",
            "cfg_dns_set_state", "cfg_dns_get_state", "cfg_dns_reset_dt", "cfg_dns_reset_time",
            "cfg_dns_reset_gradp", "cfg_dns_reset_bulkv", "cfg_dns_time", "cfg_dns_dpdx", "cfg_dns_ubulk",
            "cfg_dns_cfl", "cfg_dns_base", "cfg_dns_device_state", "cfg_dns_profile", "cfg_dns_kernel_times",
            "cfg_kernel_name", "cfg_nonlinear", "cfg_solve")
NKERNELS = 9


class CfgError(RuntimeError):
    pass


class _CFlags(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("baseflow", "constraint", "timestepping", "initstepping",
                                       "nonlinearity", "dealiasing", "taucorrection", "reserved")] + \
               [(n, C.c_double) for n in ("nu", "Vsuck", "rotation", "dPdx", "dPdz", "Ubulk", "Wbulk",
                                          "ulowerwall", "uupperwall", "wlowerwall", "wupperwall", "t0", "dt")]


class _CGrid(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("reserved", C.c_int),
                ("Lx", C.c_double), ("Lz", C.c_double), ("a", C.c_double), ("b", C.c_double)]


_lib = None


def library():
    """Load libcfgpu.so (raises CfgError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CfgError(f"{LIB_PATH} is missing: build it with `python -m channelflow_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(C.c_double)
    i, d, v = C.c_int, C.c_double, C.c_void_p
    G, F = C.POINTER(_CGrid), C.POINTER(_CFlags)
    sig = {
        "cfg_last_error": ([], C.c_char_p),
        "cfg_device_count": ([C.POINTER(C.c_int)], i),
        "cfg_dns_create": ([G, F, P, P, P, P, C.POINTER(v)], i),
        "cfg_dns_destroy": ([v], i),
        "cfg_dns_advance": ([v, i], i),
        "cfg_dns_set_state": ([v, P, P], i),
        "cfg_dns_get_state": ([v, P, P], i),
        "cfg_dns_reset_dt": ([v, d], i),
        "cfg_dns_reset_time": ([v, d], i),
        "cfg_dns_reset_gradp": ([v, d, d], i),
        "cfg_dns_reset_bulkv": ([v, d, d], i),
        "cfg_dns_time": ([v, P], i),
        "cfg_dns_dpdx": ([v, P], i),
        "cfg_dns_ubulk": ([v, P], i),
        "cfg_dns_cfl": ([v, P], i),
        "cfg_dns_base": ([v, P, P], i),
        "cfg_dns_device_state": ([v, C.POINTER(P), C.POINTER(P), C.POINTER(C.c_int), C.POINTER(C.c_int)], i),
        "cfg_dns_profile": ([v, i], i),
        "cfg_dns_kernel_times": ([v, P, C.POINTER(C.c_longlong)], i),
        "cfg_kernel_name": ([i], C.c_char_p),
        "cfg_nonlinear": ([G, F, P, P], i),
        "cfg_solve": ([G, F, d, P, P, P, P, P], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise CfgError(f"cfgpu error {rc}: {library().cfg_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


class DNSFlags:
    """channelflow's DNSFlags (dnsflags.h:84-140) with the same defaults."""

    def __init__(self, **kw):
        self.baseflow = "LaminarBase"
        self.constraint = "PressureGradient"
        self.timestepping = "SBDF3"
        self.initstepping = "SMRK2"
        self.nonlinearity = "Rotational"
        self.dealiasing = "DealiasXZ"
        self.taucorrection = True
        self.nu = 0.0025
        self.Vsuck = 0.0
        self.rotation = 0.0
        self.dPdx = 0.0
        self.dPdz = 0.0
        self.Ubulk = 0.0
        self.Wbulk = 0.0
        self.ulowerwall = 0.0
        self.uupperwall = 0.0
        self.wlowerwall = 0.0
        self.wupperwall = 0.0
        self.t0 = 0.0
        self.dt = 0.03125
        for k, val in kw.items():
            if not hasattr(self, k):
                raise CfgError(f"unknown DNSFlags field {k}")
            setattr(self, k, val)

    def dealias_xz(self):
        return self.dealiasing in ("DealiasXZ", "DealiasXYZ")

    def c(self):
        f = _CFlags()
        for k, table in _ENUMS.items():
            val = getattr(self, k)
            setattr(f, k, table[val] if isinstance(val, str) else int(val))
        f.taucorrection = int(bool(self.taucorrection))
        for k in ("nu", "Vsuck", "rotation", "dPdx", "dPdz", "Ubulk", "Wbulk", "ulowerwall", "uupperwall",
                  "wlowerwall", "wupperwall", "t0", "dt"):
            setattr(f, k, float(getattr(self, k)))
        return f

    def copy(self, **kw):
        d = dict(self.__dict__)
        d.update(kw)
        return DNSFlags(**d)


class FlowField:
    """Spectral FlowField(Nx, Ny, Nz, Nd, Lx, Lz, a, b) (flowfield.h:56) holding
    coefficients uhat(i, my, mx, mz) in channelflow's serial layout."""

    def __init__(self, Nx, Ny, Nz, Nd, Lx, Lz, a=-1.0, b=1.0, data=None):
        self.Nx, self.Ny, self.Nz, self.Nd = int(Nx), int(Ny), int(Nz), int(Nd)
        self.Lx, self.Lz, self.a, self.b = float(Lx), float(Lz), float(a), float(b)
        shape = (self.Nd, self.Ny, self.Nx, self.Nz // 2 + 1)
        self.data = np.zeros(shape, dtype=np.complex128) if data is None else \
            np.ascontiguousarray(data, dtype=np.complex128).reshape(shape).copy()

    def cmplx(self, mx, my, mz, i):
        return self.data[i, my, mx, mz]

    def kx(self, mx):
        return mx if mx <= self.Nx // 2 else mx - self.Nx

    def grid(self):
        return _CGrid(self.Nx, self.Ny, self.Nz, 0, self.Lx, self.Lz, self.a, self.b)

    def congruent(self, f):
        return (self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b) == \
            (f.Nx, f.Ny, f.Nz, f.Lx, f.Lz, f.a, f.b)


class DNS:
    """channelflow's DNS (dns.h:30-97): DNS(fields, flags) or DNS(fields, [Ubase, Wbase], flags).

    advance(fields, nSteps) copies the caller's fields in, advances on the GPU
    and copies them back, like channelflow. Between calls the velocity,
    pressure and multistep history stay on the device; advance_resident(n)
    advances without any host transfer.
    """

    def __init__(self, fields, flags_or_base, flags=None):
        base = None
        if flags is None:
            flags = flags_or_base
        else:
            base = flags_or_base
        u = fields[0]
        q = fields[1] if len(fields) > 1 else None
        self.flags = flags
        self._grid = u.grid()
        self._shape_u = u.data.shape
        self._h = C.c_void_p()
        U = W = None
        if base is not None:
            U = np.ascontiguousarray(base[0], dtype=np.float64)
            W = np.ascontiguousarray(base[1], dtype=np.float64)
        cf = flags.c()
        _check(library().cfg_dns_create(C.byref(self._grid), C.byref(cf), _ptr(_f64(u.data)),
                                        _ptr(_f64(q.data)) if q is not None else None, _ptr(U), _ptr(W),
                                        C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            library().cfg_dns_destroy(h)
            self._h = None

    def advance(self, fields, nSteps=1):
        self.set_state(fields)
        self.advance_resident(nSteps)
        self.get_state(fields)

    def advance_resident(self, nSteps=1):
        _check(library().cfg_dns_advance(self._h, int(nSteps)))

    def set_state(self, fields):
        u = _f64(fields[0].data)
        q = _f64(fields[1].data) if len(fields) > 1 else None
        _check(library().cfg_dns_set_state(self._h, _ptr(u), _ptr(q)))

    def get_state(self, fields):
        u = np.zeros(self._shape_u, dtype=np.complex128)
        q = np.zeros((1,) + self._shape_u[1:], dtype=np.complex128)
        _check(library().cfg_dns_get_state(self._h, _ptr(u.view(np.float64)), _ptr(q.view(np.float64))))
        fields[0].data[...] = u
        if len(fields) > 1:
            fields[1].data[...] = q

    def state(self):
        u = np.zeros(self._shape_u, dtype=np.complex128)
        q = np.zeros((1,) + self._shape_u[1:], dtype=np.complex128)
        _check(library().cfg_dns_get_state(self._h, _ptr(u.view(np.float64)), _ptr(q.view(np.float64))))
        return u, q

    def _scalar(self, fn):
        out = C.c_double()
        _check(fn(self._h, C.byref(out)))
        return out.value

    def time(self):
        return self._scalar(library().cfg_dns_time)

    def dPdx(self):
        return self._scalar(library().cfg_dns_dpdx)

    def Ubulk(self):
        return self._scalar(library().cfg_dns_ubulk)

    def reset_dt(self, dt):
        _check(library().cfg_dns_reset_dt(self._h, float(dt)))
        self.flags = self.flags.copy(dt=dt)

    def reset_time(self, t):
        _check(library().cfg_dns_reset_time(self._h, float(t)))

    def reset_gradp(self, dPdx, dPdz=0.0):
        _check(library().cfg_dns_reset_gradp(self._h, float(dPdx), float(dPdz)))

    def reset_bulkv(self, Ubulk, Wbulk=0.0):
        _check(library().cfg_dns_reset_bulkv(self._h, float(Ubulk), float(Wbulk)))

    def profile(self, enable=True):
        """Enable (and clear) per-kernel HIP-event timing on this DNS's stream."""
        _check(library().cfg_dns_profile(self._h, int(bool(enable))))

    def kernel_times(self):
        """{kernel name: (total ms, launches)} since profile(True)."""
        tot = np.zeros(NKERNELS)
        cnt = (C.c_longlong * NKERNELS)()
        _check(library().cfg_dns_kernel_times(self._h, _ptr(tot), cnt))
        return {library().cfg_kernel_name(i).decode(): (float(tot[i]), int(cnt[i])) for i in range(NKERNELS)}

    def Ubase(self):
        U = np.zeros(self._grid.Ny)
        W = np.zeros(self._grid.Ny)
        _check(library().cfg_dns_base(self._h, _ptr(U), _ptr(W)))
        return U, W


def nonlinear(u, flags):
    """One NSE::nonlinear call on the GPU (nse.cpp:383-391): returns f (spectral)."""
    f = np.zeros_like(u.data)
    g = u.grid()
    cf = flags.c()
    _check(library().cfg_nonlinear(C.byref(g), C.byref(cf), _ptr(_f64(u.data)), _ptr(f.view(np.float64))))
    return f


def solve(u0, rhs, flags, lambda_t, u_init=None, q_init=None):
    """One NSE::solve call on the GPU with TauSolvers for lambda_t (nse.cpp:479-575)."""
    g = u0.grid()
    cf = flags.c()
    uo = np.zeros_like(u0.data) if u_init is None else np.ascontiguousarray(u_init, dtype=np.complex128).copy()
    qo = np.zeros((1,) + u0.data.shape[1:], dtype=np.complex128) if q_init is None else \
        np.ascontiguousarray(q_init, dtype=np.complex128).copy()
    dp = C.c_double()
    _check(library().cfg_solve(C.byref(g), C.byref(cf), float(lambda_t), _ptr(_f64(u0.data)), _ptr(_f64(rhs)),
                               _ptr(uo.view(np.float64)), _ptr(qo.view(np.float64)), C.byref(dp)))
    return uo, qo, dp.value


def device_count():
    n = C.c_int()
    _check(library().cfg_device_count(C.byref(n)))
    return n.value
