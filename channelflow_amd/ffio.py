"""channelflow's binary FlowField format (.ff), spectral state.

Layout (FlowField::binarySave, flowfield.cpp:2642-2723, read back by
FlowField(filebase), flowfield.cpp:215-300), everything big-endian
(cfbasics/cfbasics.h:367-470):
  int32 major, minor, update; int32 Nx, Ny, Nz, Nd;
  char xzstate, ystate ('S' spectral, 'P' physical);
  float64 Lx, Lz, a, b; char padded ('1' / '0');
  then the coefficients as (re, im) float64 pairs, component, then y index,
  then kx row, then kz:
    padded:   kx rows 0..Nxd and Nx-Nxd..Nx-1 (Nxd = 2 (Nx/6)), kz 0..Nz/3;
    unpadded: every kx row, kz 0..Nz/2.
Physical-state files are refused: this library holds spectral fields only.
"""
import os
import struct

import numpy as np

VERSION = (2, 0, 2)  # channelflow 2.0.2


def _path(filebase):
    return filebase if filebase.endswith(".ff") else filebase + ".ff"


def _rows(Nx, padded):
    if not padded:
        return list(range(Nx))
    Nxd = 2 * (Nx // 6)
    return list(range(0, Nxd + 1)) + list(range(Nx - Nxd, Nx))


def save_ff(filebase, data, Nx, Ny, Nz, Lx, Lz, a=-1.0, b=1.0, padded=True):
    """Write spectral coefficients data[i, ny, mx, mz] (Nd, Ny, Nx, Nz/2+1)."""
    data = np.asarray(data, dtype=np.complex128)
    Nd = data.shape[0]
    if data.shape != (Nd, Ny, Nx, Nz // 2 + 1):
        raise ValueError(f"save_ff: data shape {data.shape} is not (Nd, {Ny}, {Nx}, {Nz // 2 + 1})")
    nz = Nz // 3 + 1 if padded else Nz // 2 + 1
    rows = _rows(Nx, padded)
    body = data[:, :, rows, :nz]                    # (Nd, Ny, rows, nz)
    out = np.empty(body.shape + (2,), dtype=">f8")
    out[..., 0] = body.real
    out[..., 1] = body.imag
    head = struct.pack(">7i", *VERSION, Nx, Ny, Nz, Nd) + b"SS" + struct.pack(">4d", Lx, Lz, a, b)
    head += b"1" if padded else b"0"
    with open(_path(filebase), "wb") as f:
        f.write(head)
        f.write(out.tobytes())


def load_ff(filebase):
    """Read a spectral .ff file: (data, dict(Nx, Ny, Nz, Nd, Lx, Lz, a, b, padded, version))."""
    path = filebase if os.path.exists(filebase) else _path(filebase)
    with open(path, "rb") as f:
        raw = f.read()
    major, minor, update, Nx, Ny, Nz, Nd = struct.unpack(">7i", raw[:28])
    xz, ys = chr(raw[28]), chr(raw[29])
    Lx, Lz, a, b = struct.unpack(">4d", raw[30:62])
    padded = raw[62:63] != b"0"
    if xz != "S" or ys != "S":
        raise ValueError(f"load_ff: {path} holds a field in state ({xz}, {ys}); only spectral fields are read")
    nz = Nz // 3 + 1 if padded else Nz // 2 + 1
    rows = _rows(Nx, padded)
    n = Nd * Ny * len(rows) * nz * 2
    vals = np.frombuffer(raw, dtype=">f8", count=n, offset=63)
    if vals.size != n or len(raw) != 63 + 8 * n:
        raise ValueError(f"load_ff: {path} has {len(raw)} bytes, expected {63 + 8 * n}")
    body = vals.astype(np.float64).reshape(Nd, Ny, len(rows), nz, 2)
    data = np.zeros((Nd, Ny, Nx, Nz // 2 + 1), dtype=np.complex128)
    data[:, :, rows, :nz] = body[..., 0] + 1j * body[..., 1]
    meta = dict(Nx=Nx, Ny=Ny, Nz=Nz, Nd=Nd, Lx=Lx, Lz=Lz, a=a, b=b, padded=padded, version=(major, minor, update))
    return data, meta
