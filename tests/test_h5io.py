"""channelflow's HDF5 and NetCDF4 FlowField files (channelflow_amd/h5io.py over
libcfgh5.so), CPU only.

  * The reference's own NetCDF4 fixture tests/data/uinit.nc is read natively:
    the physical values equal the committed conversion (tests/golden/uinit.f64,
    made with h5dump) bit for bit, and the spectral field equals the oracle's
    de-padding of the same values (readNetCDF, flowfield.cpp:3605-3826).
  * .h5 files are written in FlowField::hdf5Save's layout (flowfield.cpp:
    2727-2782, 4294-4436): attribute and dataset names, big-endian types, the
    (Nd, Nz, Ny, Nx) dataspace over nz-fastest data, the grid points; checked
    with h5dump where it is installed.
  * save -> load round trips, padded and unpadded, as ioTest
    (tests/ioTest.cpp:51-73, bar 1e-14).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import cfref as R
from channelflow_amd import h5io
from channelflow_amd.fields import random_field

HERE = os.path.dirname(os.path.abspath(__file__))
UINIT_NC = "/root/reference/tests/data/uinit.nc"
LX, LZ = 2 * np.pi / 1.14, 2 * np.pi / 2.5

needs_lib = pytest.mark.skipif(not os.path.exists(h5io.LIB_PATH), reason="libcfgh5.so not built (no HDF5 found)")


@needs_lib
@pytest.mark.skipif(not os.path.exists(UINIT_NC), reason="the reference checkout is not mounted here")
def test_read_reference_netcdf_fixture():
    vals, info = h5io.nc_physical(UINIT_NC)
    man = json.load(open(os.path.join(HERE, "golden", "manifest.json")))["uinit"]
    ref = np.fromfile(os.path.join(HERE, "golden", "uinit.f64")).reshape(man["shape"])
    assert vals.shape == tuple(man["shape"])
    assert np.array_equal(vals, ref)  # the same bytes as the h5dump conversion
    assert (info["Nx"], info["Ny"], info["Nz"]) == (man["Nx"], man["Ny"], man["Nz"])
    assert info["Lx"] == man["Lx"] and info["Lz"] == man["Lz"] and (info["a"], info["b"]) == (man["a"], man["b"])
    u, m = h5io.load_nc(UINIT_NC)
    g = R.Grid(m["Nx"], m["Ny"], m["Nz"], m["Lx"], m["Lz"], m["a"], m["b"])
    uo = R.field_from_io(g, ref)
    kxd, kzd = m["Nx"] // 3 - 1, m["Nz"] // 3 - 1
    kx = np.fft.fftfreq(m["Nx"], 1.0 / m["Nx"])
    live = (np.abs(kx)[:, None] <= kxd) & (np.arange(m["Nz"] // 2 + 1)[None, :] <= kzd)
    err = np.max(np.abs((u - uo)[..., live])) / np.max(np.abs(uo))
    print("uinit.nc spectral field vs the oracle's de-padding:", err)
    assert err < 1e-14
    assert np.max(np.abs(u[..., ~live])) < 1e-14 * np.max(np.abs(u))


@needs_lib
@pytest.mark.parametrize("padded", [True, False])
def test_h5_round_trip(tmp_path, padded):
    Nx, Ny, Nz = 24, 17, 18
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.3, decay=0.5, seed=3, dealias=padded)
    path = h5io.save_h5(str(tmp_path / "u"), u, Nx, Ny, Nz, LX, LZ, padded=padded)
    assert path.endswith(".h5")
    info = h5io.h5_info(path)
    io = ((2 * Nx) // 3, (2 * Nz) // 3) if padded else (Nx, Nz)
    assert (info["Nx"], info["Ny"], info["Nz"], info["Nd"]) == (io[0], Ny, io[1], 3)
    assert (info["Nxpad"], info["Nypad"], info["Nzpad"]) == (Nx, Ny, Nz)
    assert (info["Lx"], info["Lz"], info["a"], info["b"]) == (LX, LZ, -1.0, 1.0)
    v, m, pad = h5io.load_h5(path)
    assert pad == padded
    err = np.max(np.abs(v - u)) / np.max(np.abs(u))
    print("padded" if padded else "unpadded", "round trip max rel diff", err)
    assert err < 1e-14


@needs_lib
def test_h5_layout_matches_reference_writer(tmp_path):
    """Physical values land at flat index nz + Nz*(ny + Ny*(nx + Nx*i)) of /data/u
    (hdf5write, flowfield.cpp:4338-4361); compared against this field's values
    from the oracle's own transform."""
    Nx, Ny, Nz = 8, 9, 12
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.2, decay=0.6, seed=8, dealias=False)
    path = h5io.save_h5(str(tmp_path / "v.h5"), u, Nx, Ny, Nz, LX, LZ, padded=False)
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    phys = R.field_physical(g, u)                          # (Nd, Ny, Nx, Nz) from the oracle
    flat = np.transpose(phys, (0, 2, 1, 3)).ravel()        # [i][nx][ny][nz]
    buf = np.zeros(flat.size)
    import ctypes as C
    h5io._check(h5io.library().cfgio_h5_read(path.encode(), (C.c_int * 4)(Nx, Ny, Nz, 3), h5io._p(buf)))
    assert np.max(np.abs(buf - flat)) < 1e-15 * np.max(np.abs(flat)) * 10
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        pytest.skip("h5dump not installed")
    head = subprocess.run([h5dump, "-H", path], capture_output=True, text=True, check=True).stdout
    for name in ("Nx", "Ny", "Nz", "Nd", "Nxpad", "Nypad", "Nzpad"):
        assert f'ATTRIBUTE "{name}"' in head
    assert head.count("H5T_STD_I32BE") == 7 and "H5T_IEEE_F64BE" in head
    assert 'DATASET "u"' in head and f"( 3, {Nz}, {Ny}, {Nx} )" in head
    for name in ("x", "y", "z"):
        assert f'DATASET "{name}"' in head
    y = subprocess.run([h5dump, "-m", "%.17g", "-d", "/geom/y", path], capture_output=True, text=True,
                       check=True).stdout
    vals = [float(t) for t in y.split("DATA {")[1].split("}")[0].replace("(0):", "").replace(",", " ").split()
            if not t.startswith("(")]
    assert np.allclose(vals, np.cos(np.pi * np.arange(Ny) / (Ny - 1)), rtol=0, atol=1e-16)


@needs_lib
def test_flowfield_save_and_load_h5(tmp_path):
    import channelflow_amd as cfa
    Nx, Ny, Nz = 16, 17, 16
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.1, decay=0.4, seed=12)
    f = cfa.FlowField(Nx, Ny, Nz, 3, LX, LZ, data=u)
    f.save(str(tmp_path / "w.h5"))
    h = cfa.FlowField.from_file(str(tmp_path / "w.h5"))
    assert (h.Nx, h.Ny, h.Nz, h.Nd) == (Nx, Ny, Nz, 3) and (h.Lx, h.Lz) == (LX, LZ)
    assert np.max(np.abs(h.data - u)) < 1e-14 * np.max(np.abs(u))
