"""channelflow's binary .ff format (FlowField::binarySave / FlowField(filebase),
flowfield.cpp:2642-2723, 215-300): round trips, and the exact bytes of a small
field against the layout assembled value by value (big-endian ints and doubles,
'S'/'P' and '1'/'0' characters, component / y / kx row / kz order)."""
import struct

import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.ffio import load_ff, save_ff
from channelflow_amd.fields import random_field

LX, LZ = 2 * np.pi / 1.14, 2 * np.pi / 2.5


def test_roundtrip_padded_and_unpadded(tmp_path):
    u = random_field(24, 17, 18, LX, LZ, norm=0.1, decay=0.5, seed=3)
    f = cfa.FlowField(24, 17, 18, 3, LX, LZ, data=u)
    f.save(str(tmp_path / "u"))
    g = cfa.FlowField.from_file(str(tmp_path / "u"))
    assert (g.Nx, g.Ny, g.Nz, g.Nd, g.Lx, g.Lz, g.a, g.b) == (24, 17, 18, 3, LX, LZ, -1.0, 1.0)
    assert np.array_equal(g.data, f.data)      # a dealiased field fits the padded rows exactly
    v = random_field(16, 9, 16, LX, LZ, norm=0.1, decay=0.5, seed=4, dealias=False)
    save_ff(str(tmp_path / "v.ff"), v, 16, 9, 16, LX, LZ, padded=False)
    w, meta = load_ff(str(tmp_path / "v.ff"))
    assert not meta["padded"] and np.array_equal(w, v)


def test_exact_bytes(tmp_path):
    Nx, Ny, Nz = 6, 3, 6
    rng = np.random.default_rng(0)
    d = rng.standard_normal((1, Ny, Nx, Nz // 2 + 1)) + 1j * rng.standard_normal((1, Ny, Nx, Nz // 2 + 1))
    save_ff(str(tmp_path / "t"), d, Nx, Ny, Nz, 2.0, 3.0, -1.0, 1.0, padded=True)
    exp = b"".join(struct.pack(">i", v) for v in (2, 0, 2, Nx, Ny, Nz, 1))
    exp += b"S" + b"S" + b"".join(struct.pack(">d", v) for v in (2.0, 3.0, -1.0, 1.0)) + b"1"
    Nxd = 2 * (Nx // 6)
    for ny in range(Ny):
        for nx in list(range(0, Nxd + 1)) + list(range(Nx - Nxd, Nx)):
            for nz in range(0, (2 * (Nz // 3) + 1) // 2 + 1):
                exp += struct.pack(">d", d[0, ny, nx, nz].real) + struct.pack(">d", d[0, ny, nx, nz].imag)
    assert (tmp_path / "t.ff").read_bytes() == exp


def test_physical_state_refused(tmp_path):
    p = tmp_path / "p.ff"
    p.write_bytes(struct.pack(">7i", 2, 0, 2, 4, 3, 4, 1) + b"PP" + struct.pack(">4d", 1, 1, -1, 1) + b"0")
    with pytest.raises(ValueError):
        load_ff(str(p))
