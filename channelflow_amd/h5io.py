"""channelflow's HDF5 (.h5) and NetCDF4 (.nc) FlowField files.

The files hold the physical velocity on the de-padded grid; the spectral
field of a padded FlowField is recovered as channelflow does it:

  * save (FlowField::hdf5Save, flowfield.cpp:2727-2782): a padded field
    (|kx| > Nx/3 - 1 and kz > Nz/3 - 1 zero) is first interpolated onto the
    (2 Nx/3) x Ny x (2 Nz/3) grid (FlowField::interpolate, flowfield.cpp:
    691-794), then transformed to physical values and written with the grid
    points and attributes (layout in csrc/cfg_h5io.cpp);
  * load (FlowField(filebase), flowfield.cpp:104-215 for .h5, readNetCDF
    :3605-3826 for .nc): the values are transformed to spectral on the file's
    grid and interpolated onto the padded grid of the attributes.

Transforms follow FlowField::makeSpectral / makePhysical (flowfield.cpp:
1850-1987): an xz real FFT scaled by 1/(Nx Nz), and in y the DCT-I of the
values at y_n = cos(pi n/(Ny-1)) scaled by 1/(Ny-1) with the end coefficients
halved. The HDF5 reads and writes go through libcfgh5.so (the HDF5 C library);
these conversions run on the host, they are not part of the time step.

Spectral arrays have channelflow's shape (Nd, Ny, Nx, Nz/2+1).
"""
import ctypes as C
import os

import numpy as np
from scipy.fft import dct

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfgh5.so")
_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m channelflow_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.cfgio_last_error.restype = C.c_char_p
        L.cfgio_h5_save.argtypes = [C.c_char_p, I, P, P, P, P, P]
        L.cfgio_h5_info.argtypes = [C.c_char_p, I, P]
        L.cfgio_h5_read.argtypes = [C.c_char_p, I, P]
        L.cfgio_nc_info.argtypes = [C.c_char_p, I, P]
        L.cfgio_nc_read.argtypes = [C.c_char_p, C.c_int, P, C.c_long]
        for f in ("cfgio_h5_save", "cfgio_h5_info", "cfgio_h5_read", "cfgio_nc_info", "cfgio_nc_read"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise IOError("cfgio: " + library().cfgio_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# transforms (FlowField::makeSpectral / makePhysical)
# ---------------------------------------------------------------------------
def spectral_from_physical(phys):
    """phys (Nd, Ny, Nx, Nz) real values -> (Nd, Ny, Nx, Nz/2+1) coefficients."""
    nd, Ny, Nx, Nz = phys.shape
    xz = np.fft.rfft2(phys, axes=(2, 3)) / (Nx * Nz)
    c = dct(xz, type=1, axis=1) / (Ny - 1)
    c[:, 0] *= 0.5
    c[:, -1] *= 0.5
    return c


def physical_from_spectral(spec, Nz):
    """(Nd, Ny, Nx, Nz/2+1) coefficients -> (Nd, Ny, Nx, Nz) real values."""
    nd, Ny, Nx, Mz = spec.shape
    c = np.array(spec, dtype=np.complex128)
    c[:, 1:-1] *= 0.5
    y = dct(c, type=1, axis=1)
    return np.fft.irfft2(y, s=(Nx, Nz), axes=(2, 3)) * (Nx * Nz)


def interpolate(f, Nzf, Nxg, Nzg):
    """FlowField::interpolate (flowfield.cpp:691-794) in x and z, same Ny: the
    modes -Nx/2+1 <= kx < Nx/2 and 0 <= kz <= Nz/2 common to both grids are
    copied; going to more x modes, f's kx = Nxf/2 column lands on g's kx = -Nxf/2."""
    nd, Ny, Nxf, Mzf = f.shape
    g = np.zeros((nd, Ny, Nxg, Nzg // 2 + 1), dtype=np.complex128)
    kxmax = min(Nxf // 2, Nxg // 2)
    kxmin = max(-Nxf // 2 + 1, -Nxg // 2 + 1)
    kzmax = min(Nzf // 2, Nzg // 2)
    for kx in range(kxmin, kxmax):
        g[:, :, kx % Nxg, :kzmax + 1] = f[:, :, kx % Nxf, :kzmax + 1]
    if Nxg > Nxf:
        kx = Nxf // 2
        g[:, :, (-kx) % Nxg, :Nzf // 2 + 1] = f[:, :, kx % Nxf, :Nzf // 2 + 1]
    return g


def _grid_points(Nx, Ny, Nz, Lx, Lz, a, b):
    x = np.arange(Nx) * (Lx / Nx)
    z = np.arange(Nz) * (Lz / Nz)
    y = (b + a) / 2 + (b - a) / 2 * np.cos(np.arange(Ny) * (np.pi / (Ny - 1)))
    return x, y, z


# ---------------------------------------------------------------------------
# .h5
# ---------------------------------------------------------------------------
def save_h5(path, spec, Nx, Ny, Nz, Lx, Lz, a=-1.0, b=1.0, padded=True):
    """FlowField::hdf5Save: writes `path` (".h5" appended if missing)."""
    if not path.endswith(".h5"):
        path += ".h5"
    spec = np.asarray(spec, dtype=np.complex128)
    nd = spec.shape[0]
    Nxio, Nzio = ((2 * Nx) // 3, (2 * Nz) // 3) if padded else (Nx, Nz)
    v = interpolate(spec, Nz, Nxio, Nzio)
    phys = physical_from_spectral(v, Nzio)                              # (Nd, Ny, Nx, Nz)
    data = np.ascontiguousarray(np.transpose(phys, (0, 2, 1, 3)))        # [i][nx][ny][nz]
    x, y, z = _grid_points(Nxio, Ny, Nzio, Lx, Lz, a, b)
    ints = (C.c_int * 7)(Nxio, Ny, Nzio, nd, Nx, Ny, Nz)
    dbl = np.array([Lx, Lz, a, b], dtype=np.float64)
    _check(library().cfgio_h5_save(path.encode(), ints, _p(dbl), _p(data), _p(x), _p(y), _p(z)))
    return path


def h5_info(path):
    ints = (C.c_int * 7)()
    dbl = np.zeros(4)
    _check(library().cfgio_h5_info(path.encode(), ints, _p(dbl)))
    keys = ("Nx", "Ny", "Nz", "Nd", "Nxpad", "Nypad", "Nzpad")
    out = {k: int(v) for k, v in zip(keys, ints)}
    out.update(Lx=float(dbl[0]), Lz=float(dbl[1]), a=float(dbl[2]), b=float(dbl[3]))
    return out


def load_h5(path):
    """FlowField(filebase) from an .h5 file: (spectral array on the padded grid,
    info dict, padded flag)."""
    m = h5_info(path)
    Nx, Ny, Nz, nd = m["Nx"], m["Ny"], m["Nz"], m["Nd"]
    if m["Nypad"] != Ny:
        raise IOError("FlowField(h5): Ny != Nypad in the file")
    data = np.zeros(nd * Nx * Ny * Nz)
    _check(library().cfgio_h5_read(path.encode(), (C.c_int * 4)(Nx, Ny, Nz, nd), _p(data)))
    phys = np.transpose(data.reshape(nd, Nx, Ny, Nz), (0, 2, 1, 3))
    v = spectral_from_physical(phys)
    u = interpolate(v, Nz, m["Nxpad"], m["Nzpad"])
    same = (Nx, Nz) == (m["Nxpad"], m["Nzpad"])
    padded = (not same) and (2 * m["Nxpad"]) // 3 >= Nx and (2 * m["Nzpad"]) // 3 >= Nz
    return u, m, padded


# ---------------------------------------------------------------------------
# .nc (read)
# ---------------------------------------------------------------------------
def nc_physical(path):
    """The de-padded physical velocity of a channelflow NetCDF4 file:
    (array (3, Nz_io, Ny, Nx_io), info dict)."""
    ints = (C.c_int * 3)()
    dbl = np.zeros(4)
    _check(library().cfgio_nc_info(path.encode(), ints, _p(dbl)))
    Nx, Ny, Nz = (int(v) for v in ints)
    info = dict(Nx=Nx, Ny=Ny, Nz=Nz, Lx=float(dbl[0]), Lz=float(dbl[1]), a=float(dbl[2]), b=float(dbl[3]))
    Nxio, Nzio = 2 * (Nx // 3), 2 * (Nz // 3)
    out = np.zeros((3, Nzio, Ny, Nxio))
    for c in range(3):
        buf = np.zeros(Nzio * Ny * Nxio)
        _check(library().cfgio_nc_read(path.encode(), c, _p(buf), buf.size))
        out[c] = buf.reshape(Nzio, Ny, Nxio)
    return out, info


def load_nc(path):
    """FlowField(filebase) from a NetCDF4 file (readNetCDF): the spectral field
    on the padded grid Nx x Ny x Nz of its attributes, and the info dict."""
    vals, m = nc_physical(path)
    phys = np.transpose(vals, (0, 2, 3, 1))                              # (3, Ny, Nx_io, Nz_io)
    v = spectral_from_physical(phys)
    return interpolate(v, phys.shape[3], m["Nx"], m["Nz"]), m
