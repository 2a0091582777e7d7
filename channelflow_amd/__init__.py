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
            "cfg_dns_cfl", "cfg_dns_base", "cfg_dns_device_state", "cfg_dns_profile", "cfg_dns_set_graphs",
            "cfg_dns_kernel_times",
            "cfg_kernel_name", "cfg_nonlinear", "cfg_nonlinear_noalloc", "cfg_solve", "cfg_pressure",
            "cfg_batch_create", "cfg_batch_destroy", "cfg_batch_add", "cfg_batch_remove", "cfg_batch_advance",
            "cfg_batch_size", "cfg_batch_get_state", "cfg_batch_set_state", "cfg_batch_time", "cfg_batch_finite",
            "cfg_batch_dpdx",
            "cfg_batch_cfl", "cfg_batch_reset_dt", "cfg_batch_advance_members", "cfg_batch_profile", "cfg_batch_kernel_times", "cfg_batch_set_graphs", "cfg_batch_graph_stats", "cfg_batch_tick_stats",
            "cfg_server_create", "cfg_server_destroy", "cfg_server_submit", "cfg_server_step", "cfg_server_result",
            "cfg_server_memory", "cfg_tausolver_create", "cfg_tausolver_destroy", "cfg_tausolver_solve",
            "cfg_tausolver_solve_bulk", "cfg_tausolver_verify", "cfg_tausolver_verify_bulk")
NKERNELS = 9


class CfgError(RuntimeError):
    """A failed library call; `code` is the status (CFG_ERR_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


ERR_ARG, ERR_DEVICE, ERR_NONFINITE, ERR_STATE = -1, -2, -3, -4


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
        "cfg_dns_set_graphs": ([v, i], i),
        "cfg_dns_kernel_times": ([v, P, C.POINTER(C.c_longlong)], i),
        "cfg_kernel_name": ([i], C.c_char_p),
        "cfg_nonlinear": ([G, F, P, P], i),
        "cfg_nonlinear_noalloc": ([G, F, P, P, v], i),
        "cfg_solve": ([G, F, d, P, P, P, P, P], i),
        "cfg_batch_create": ([C.POINTER(v)], i),
        "cfg_batch_destroy": ([v], i),
        "cfg_pressure": ([G, F, P, P, P, P], i),
        "cfg_batch_add": ([v, G, F, P, P, C.POINTER(C.c_int)], i),
        "cfg_batch_remove": ([v, i], i),
        "cfg_batch_advance": ([v, i], i),
        "cfg_batch_size": ([v, C.POINTER(C.c_int), C.POINTER(C.c_int)], i),
        "cfg_batch_get_state": ([v, i, P, P], i),
        "cfg_batch_set_state": ([v, i, P, P], i),
        "cfg_batch_time": ([v, i, P], i),
        "cfg_batch_finite": ([v, i, C.POINTER(C.c_int)], i),
        "cfg_batch_dpdx": ([v, i, P], i),
        "cfg_batch_cfl": ([v, i, P], i),
        "cfg_batch_reset_dt": ([v, i, d], i),
        "cfg_batch_advance_members": ([v, i, C.POINTER(C.c_int), C.POINTER(C.c_int)], i),
        "cfg_batch_profile": ([v, i], i),
        "cfg_batch_kernel_times": ([v, P, C.POINTER(C.c_longlong)], i),
        "cfg_batch_set_graphs": ([v, i], i),
        "cfg_batch_graph_stats": ([v, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], i),
        "cfg_batch_tick_stats": ([v, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_double)], i),
        "cfg_server_create": ([d, C.POINTER(v)], i),
        "cfg_server_destroy": ([v], i),
        "cfg_server_submit": ([v, G, F, P, P, i, C.POINTER(C.c_longlong)], i),
        "cfg_server_step": ([v, C.POINTER(C.c_longlong), i, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int)], i),
        "cfg_server_result": ([v, C.c_longlong, P, P, P], i),
        "cfg_server_memory": ([v, P], i),
        "cfg_tausolver_create": ([i, i, d, d, d, d, d, d, i, i, C.POINTER(v)], i),
        "cfg_tausolver_destroy": ([v], i),
        "cfg_tausolver_solve": ([v, P, P, P, P, P, P, P], i),
        "cfg_tausolver_solve_bulk": ([v, P, P, P, d, d, P, P, P, P, P, P], i),
        "cfg_tausolver_verify": ([v, P, P, P, P, P, P, P, P], i),
        "cfg_tausolver_verify_bulk": ([v, P, P, P, P, d, d, P, P, P, d, d, P], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise CfgError(f"cfgpu error {rc}: {library().cfg_last_error().decode()}", rc)


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
        # time-step control, used on the host by TimeStep (dnsflags.h:86-87, 129-136)
        self.T = 20.0
        self.dT = 1.0
        self.variabledt = True
        self.dtmin = 0.001
        self.dtmax = 0.2
        self.CFLmin = 0.4
        self.CFLmax = 0.6
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

    def save(self, filebase, padded=True):
        """FlowField::save (flowfield.cpp:2598-2640): ".h5" writes channelflow's
        HDF5 layout (hdf5Save, de-padded physical values), anything else the
        binary .ff format. NetCDF files are read (from_file), not written."""
        if filebase.endswith(".nc"):
            raise CfgError("FlowField.save: NetCDF files are read, not written, by this library; use .h5 or .ff")
        if filebase.endswith(".h5"):
            from .h5io import save_h5
            save_h5(filebase, self.data, self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b, padded)
            return
        from .ffio import save_ff
        save_ff(filebase, self.data, self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b, padded)

    @classmethod
    def from_file(cls, filebase):
        """FlowField(filebase) (flowfield.cpp:104-300): a NetCDF4 .nc, an HDF5
        .h5 or a binary .ff file, tried in that order when no suffix is given."""
        cands = [filebase] if filebase.endswith((".nc", ".h5", ".ff")) else \
            [filebase + s for s in (".nc", ".h5", ".ff")]
        for path in cands:
            if not os.path.exists(path):
                continue
            if path.endswith(".nc"):
                from .h5io import load_nc
                data, m = load_nc(path)
                return cls(m["Nx"], m["Ny"], m["Nz"], 3, m["Lx"], m["Lz"], m["a"], m["b"], data=data)
            if path.endswith(".h5"):
                from .h5io import load_h5
                data, m, _ = load_h5(path)
                return cls(m["Nxpad"], m["Nypad"], m["Nzpad"], m["Nd"], m["Lx"], m["Lz"], m["a"], m["b"], data=data)
            from .ffio import load_ff
            data, m = load_ff(path)
            return cls(m["Nx"], m["Ny"], m["Nz"], m["Nd"], m["Lx"], m["Lz"], m["a"], m["b"], data=data)
        raise CfgError(f"FlowField({filebase}): no readable .nc, .h5 or .ff file")

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
        self.flags = self.flags.copy(dt=float(dt))

    def CFL(self, u=None):
        """DNS::CFL(u) (dns.cpp:282-289, dnsalgo.cpp:63-67) of the velocity this
        DNS holds. u, if given, must be that velocity (fields[0] after advance,
        as every reference caller passes, simulateflow.cpp:140-142); it is
        accepted for the reference's signature, not read."""
        if u is not None and tuple(u.data.shape) != tuple(self._shape_u):
            raise CfgError("CFL(u): u is not this DNS's velocity field")
        return self._scalar(library().cfg_dns_cfl)

    def reset_time(self, t):
        _check(library().cfg_dns_reset_time(self._h, float(t)))

    def reset_gradp(self, dPdx, dPdz=0.0):
        _check(library().cfg_dns_reset_gradp(self._h, float(dPdx), float(dPdz)))

    def reset_bulkv(self, Ubulk, Wbulk=0.0):
        _check(library().cfg_dns_reset_bulkv(self._h, float(Ubulk), float(Wbulk)))

    def profile(self, enable=True):
        """Enable (and clear) per-kernel HIP-event timing on this DNS's stream."""
        _check(library().cfg_dns_profile(self._h, int(bool(enable))))

    def set_graphs(self, enable=True):
        """hipGraph replay of repeating steps (off by default: for one field it
        is no faster than eager launches; bit-identical to them)."""
        _check(library().cfg_dns_set_graphs(self._h, int(bool(enable))))

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


class TauSolver:
    """channelflow's per-mode TauSolver (tausolver.h:33-66) on the device:
    TauSolver(kx, kz, Lx, Lz, a, b, lambda, nu, Ny, tauCorrection=True).

    solve(Rx, Ry, Rz) -> (u, v, w, P) for complex Chebyshev coefficient arrays of
    length Ny; solve_bulk(Rx, Ry, Rz, umean, wmean) -> (u, v, w, P, dPdx, dPdz)
    for kx = kz = 0; verify(u, v, w, P, Rx, Ry, Rz) and verify_bulk(...) return
    the reference's residual sum."""

    def __init__(self, kx, kz, Lx, Lz, a, b, lam, nu, Ny, tauCorrection=True):
        self.Ny = int(Ny)
        self.kx, self.kz = int(kx), int(kz)
        self.lam, self.nu = float(lam), float(nu)
        h = C.c_void_p()
        _check(library().cfg_tausolver_create(self.kx, self.kz, float(Lx), float(Lz), float(a), float(b), self.lam,
                                              self.nu, self.Ny, int(bool(tauCorrection)), C.byref(h)))
        self._h = h

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape != (self.Ny,):
            raise ValueError(f"expected {self.Ny} Chebyshev coefficients, got shape {x.shape}")
        return x.view(np.float64)

    def _outs(self):
        return [np.zeros(self.Ny, dtype=np.complex128) for _ in range(4)]

    def solve(self, Rx, Ry, Rz):
        R = [self._vec(r) for r in (Rx, Ry, Rz)]
        o = self._outs()
        _check(library().cfg_tausolver_solve(self._h, *[_ptr(r) for r in R], *[_ptr(x.view(np.float64)) for x in o]))
        return tuple(o)

    def solve_bulk(self, Rx, Ry, Rz, umean, wmean):
        R = [self._vec(r) for r in (Rx, Ry, Rz)]
        o = self._outs()
        dx, dz = C.c_double(), C.c_double()
        _check(library().cfg_tausolver_solve_bulk(self._h, *[_ptr(r) for r in R], float(umean), float(wmean),
                                                  *[_ptr(x.view(np.float64)) for x in o], C.byref(dx), C.byref(dz)))
        return (*o, dx.value, dz.value)

    def verify(self, u, v, w, P, Rx, Ry, Rz):
        X = [self._vec(x) for x in (u, v, w, P, Rx, Ry, Rz)]
        e = C.c_double()
        _check(library().cfg_tausolver_verify(self._h, *[_ptr(x) for x in X], C.byref(e)))
        return e.value

    def verify_bulk(self, u, v, w, P, dPdx, dPdz, Rx, Ry, Rz, umean, wmean):
        U = [self._vec(x) for x in (u, v, w, P)]
        R = [self._vec(x) for x in (Rx, Ry, Rz)]
        e = C.c_double()
        _check(library().cfg_tausolver_verify_bulk(self._h, *[_ptr(x) for x in U], float(dPdx), float(dPdz),
                                                   *[_ptr(x) for x in R], float(umean), float(wmean), C.byref(e)))
        return e.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            library().cfg_tausolver_destroy(self._h)
            self._h = None


class Batch:
    """Many independent DNS integrations advanced together on one GPU.

    add(fields, flags) takes the same arguments as DNS(fields, flags) and
    returns a member id; advance(n) is DNS::advance(fields, n) for every member
    at once. Members on the same grid share batched launches; each member's
    result is the same as that of its own DNS.
    """

    def __init__(self):
        self._h = C.c_void_p()
        _check(library().cfg_batch_create(C.byref(self._h)))
        self._shapes = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            library().cfg_batch_destroy(h)
            self._h = None

    def add(self, fields, flags):
        u = fields[0]
        q = fields[1] if len(fields) > 1 else None
        g = u.grid()
        cf = flags.c()
        mid = C.c_int()
        _check(library().cfg_batch_add(self._h, C.byref(g), C.byref(cf), _ptr(_f64(u.data)),
                                       _ptr(_f64(q.data)) if q is not None else None, C.byref(mid)))
        self._shapes[mid.value] = u.data.shape
        return mid.value

    def remove(self, member):
        _check(library().cfg_batch_remove(self._h, int(member)))
        self._shapes.pop(member, None)

    def advance(self, nSteps=1):
        _check(library().cfg_batch_advance(self._h, int(nSteps)))

    def size(self):
        n, k = C.c_int(), C.c_int()
        _check(library().cfg_batch_size(self._h, C.byref(n), C.byref(k)))
        return n.value, k.value

    def state(self, member):
        shape = self._shapes[member]
        u = np.zeros(shape, dtype=np.complex128)
        q = np.zeros((1,) + shape[1:], dtype=np.complex128)
        _check(library().cfg_batch_get_state(self._h, int(member), _ptr(u.view(np.float64)),
                                             _ptr(q.view(np.float64))))
        return u, q

    def set_state(self, member, fields):
        u = _f64(fields[0].data)
        q = _f64(fields[1].data) if len(fields) > 1 else None
        _check(library().cfg_batch_set_state(self._h, int(member), _ptr(u), _ptr(q)))

    def time(self, member):
        t = C.c_double()
        _check(library().cfg_batch_time(self._h, int(member), C.byref(t)))
        return t.value

    def finite(self, member):
        """Whether the member's velocity was finite after its last advance (a
        CFG_ERR_NONFINITE of advance names the members for which it was not)."""
        v = C.c_int()
        _check(library().cfg_batch_finite(self._h, int(member), C.byref(v)))
        return bool(v.value)

    def dPdx(self, member):
        v = C.c_double()
        _check(library().cfg_batch_dpdx(self._h, int(member), C.byref(v)))
        return v.value

    def CFL(self, member):
        """DNS::CFL of the member's current velocity."""
        v = C.c_double()
        _check(library().cfg_batch_cfl(self._h, int(member), C.byref(v)))
        return v.value

    def reset_dt(self, member, dt):
        """DNS::reset_dt(dt) of one member (dns.cpp:180-185)."""
        _check(library().cfg_batch_reset_dt(self._h, int(member), float(dt)))

    def advance_members(self, nsteps):
        """DNS::advance(fields, n) of each member in {member: n}, together; the
        other members do not move."""
        ids = list(nsteps)
        mem = (C.c_int * max(1, len(ids)))(*ids)
        cnt = (C.c_int * max(1, len(ids)))(*[int(nsteps[k]) for k in ids])
        _check(library().cfg_batch_advance_members(self._h, len(ids), mem, cnt))

    def profile(self, enable=True):
        _check(library().cfg_batch_profile(self._h, int(bool(enable))))

    def kernel_times(self):
        tot = np.zeros(NKERNELS)
        cnt = (C.c_longlong * NKERNELS)()
        _check(library().cfg_batch_kernel_times(self._h, _ptr(tot), cnt))
        return {library().cfg_kernel_name(i).decode(): (float(tot[i]), int(cnt[i])) for i in range(NKERNELS)}

    GRAPHS_OFF, GRAPHS_PER_CLASS, GRAPHS_WHOLE = 0, 1, 2

    def set_graphs(self, mode=True):
        """hipGraph capture and replay of repeating ticks: False / 0 eager
        launches, True / 1 one graph per grid class and tick (the default),
        2 one graph per tick for the whole batch."""
        m = self.GRAPHS_PER_CLASS if mode is True else (self.GRAPHS_OFF if mode is False else int(mode))
        _check(library().cfg_batch_set_graphs(self._h, m))

    def graph_stats(self):
        """(captures, replays), whole-batch and per-class graphs together."""
        c, r = C.c_longlong(), C.c_longlong()
        _check(library().cfg_batch_graph_stats(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def tick_stats(self):
        """(ticks, graph launches, host microseconds spent in the ticks)."""
        t, g, h = C.c_longlong(), C.c_longlong(), C.c_double()
        _check(library().cfg_batch_tick_stats(self._h, C.byref(t), C.byref(g), C.byref(h)))
        return t.value, g.value, h.value


class PressureSolver:
    """PressureSolver(u, Ubase, Wbase, nu, Vsuck, nonlinearity) (poissonsolver.h:63-97),
    on the device. solve(q, u) fills q with the pressure of u: lapl q = -div N(u+U),
    dq/dy = nu v'' at the walls (the initial q of simulateflow.cpp:104-105). With
    Ubase = Wbase = None the base flow is the one a DNS with `flags` builds."""

    def __init__(self, u, Ubase=None, Wbase=None, nu=0.0025, Vsuck=0.0, nonlinearity="Rotational", flags=None):
        self.grid = u.grid()
        f = flags if flags is not None else DNSFlags()
        self.flags = f.copy(nu=nu, Vsuck=Vsuck, nonlinearity=nonlinearity) if flags is None else f
        Ny = self.grid.Ny
        self.U = None if Ubase is None else np.ascontiguousarray(np.asarray(Ubase, dtype=np.float64)[:Ny])
        self.W = None if Wbase is None else np.ascontiguousarray(np.asarray(Wbase, dtype=np.float64)[:Ny])
        if self.U is not None and self.W is None:
            self.W = np.zeros(Ny)

    def solve(self, q, u):
        uu = _f64(u.data)
        out = np.zeros(q.data.shape, dtype=np.complex128)
        _check(library().cfg_pressure(C.byref(self.grid), C.byref(self.flags.c()), _ptr(self.U), _ptr(self.W),
                                      _ptr(uu), _ptr(out.view(np.float64))))
        q.data[...] = out
        return q


class Server:
    """Continuous batching of DNS integration requests on one GPU.

    submit(fields, flags, nsteps) queues a request and returns its id; step()
    admits queued requests while their device memory fits the budget, runs
    one tick of every running request and returns the ids whose results have
    reached host memory; result(id) returns (u, q, t) and releases it.
    """

    def __init__(self, budget_bytes):
        self._h = C.c_void_p()
        _check(library().cfg_server_create(float(budget_bytes), C.byref(self._h)))
        self._shapes = {}
        self._done = (C.c_longlong * 256)()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            library().cfg_server_destroy(h)
            self._h = None

    def submit(self, fields, flags, nsteps):
        u = fields[0]
        q = fields[1] if len(fields) > 1 else None
        g = u.grid()
        cf = flags.c()
        rid = C.c_longlong()
        _check(library().cfg_server_submit(self._h, C.byref(g), C.byref(cf), _ptr(_f64(u.data)),
                                           _ptr(_f64(q.data)) if q is not None else None, int(nsteps),
                                           C.byref(rid)))
        self._shapes[rid.value] = u.data.shape
        return rid.value

    def step(self):
        """One server step: returns (completed ids, running, queued)."""
        nd, nr, nq = C.c_int(), C.c_int(), C.c_int()
        _check(library().cfg_server_step(self._h, self._done, len(self._done), C.byref(nd), C.byref(nr),
                                         C.byref(nq)))
        return [int(self._done[i]) for i in range(nd.value)], nr.value, nq.value

    def result(self, rid):
        shape = self._shapes.pop(rid)
        u = np.zeros(shape, dtype=np.complex128)
        q = np.zeros((1,) + shape[1:], dtype=np.complex128)
        t = C.c_double()
        _check(library().cfg_server_result(self._h, C.c_longlong(rid), _ptr(u.view(np.float64)),
                                           _ptr(q.view(np.float64)), C.byref(t)))
        return u, q, t.value

    def memory(self):
        v = C.c_double()
        _check(library().cfg_server_memory(self._h, C.byref(v)))
        return v.value


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


from .timestep import TimeStep  # noqa: E402  (host-side dt control)
