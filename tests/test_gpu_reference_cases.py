"""The reference's own DNS tests run on the HIP path through the C ABI, each
against its reference bar and against the oracle.

  * dnsOrrsommTest (tests/dnsOrrsommTest.cpp:39-291): the Orr-Sommerfeld mode
    at Re = 7500, every form, base or fluctuation, bulk velocity or pressure
    gradient. The device's state after 20 steps equals the oracle's within
    1e-10 per velocity component, and the device's full run (650 steps)
    meets the reference's bar on its own.
  * dnsParabolaTest (tests/dnsParabolaTest.cpp:60-248), every registered
    variant: the parabola stays steady on the device (sum of 10 samples
    < 1e-13), with the driving gradient -2 nu.
  * dnsZeroTest (tests/dnsZeroTest.cpp:63-224), every registered variant.
"""
import numpy as np
import pytest

import cfref as R
import channelflow_amd as cfa
from oscase import OSCase, NX, NZ
from test_oracle import SCHEMES, ZERO_MAXERR, OS_FORMS

pytestmark = pytest.mark.gpu


def _gpu_dns(g, flags, u, q, base=None):
    fu = cfa.FlowField(g.Nx, g.Ny, g.Nz, 3, g.Lx, g.Lz, g.a, g.b, data=u)
    fq = cfa.FlowField(g.Nx, g.Ny, g.Nz, 1, g.Lx, g.Lz, g.a, g.b, data=q)
    return cfa.DNS([fu, fq], base, flags) if base is not None else cfa.DNS([fu, fq], flags)


def _rel(g, a, b, padded):
    out = []
    for i in range(a.shape[0]):
        d = np.sqrt(R.l2dist2(g, a[i:i + 1], b[i:i + 1], padded=padded))
        n = np.sqrt(R.l2norm2(g, b[i:i + 1], padded=padded))
        out.append(d / n if n > 0 else d)
    return np.array(out)


@pytest.mark.parametrize("nonlinearity", OS_FORMS)
@pytest.mark.parametrize("base", [True, False], ids=["base", "fluc"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_gpu_orr_sommerfeld(gpu, nonlinearity, base, constraint):
    c = OSCase(base=base)
    u0, q0 = c.initial(nonlinearity)
    fl = c.flags(nonlinearity, constraint)
    print("Orr-Sommerfeld 4x65x4", nonlinearity, "base" if base else "fluc", constraint, flush=True)
    d = _gpu_dns(c.g, cfa.DNSFlags(**fl), u0, q0, (c.Ubase, c.Wbase) if base else None)
    o = R.OracleDNS(c.g, R.make_flags(**fl), u0, q0, padded=False, base=(c.Ubase, c.Wbase) if base else None)
    d.advance_resident(20)
    o.advance(20)
    ug, qg = d.state()
    uo, qo = o.state()
    # the device holds live modes only; the oracle's q keeps the initial q's
    # Nyquist modes, which no solve writes (nse.cpp:497-499)
    qo[:, :, NX // 2, :] = 0.0
    qo[:, :, :, NZ // 2] = 0.0
    err = _rel(c.g, ug, uo, padded=False)
    errq = _rel(c.g, qg, qo, padded=False)
    print("  20 steps: rel L2 per component", err, "q", errq, flush=True)
    assert np.all(err <= 1e-10), err
    assert np.all(errq <= 1e-10), errq
    # the reference's measure on the device's own run
    d2 = _gpu_dns(c.g, cfa.DNSFlags(**fl), u0, q0, (c.Ubase, c.Wbase) if base else None)
    if nonlinearity == "Rotational":
        full = c.dist

        def no_nyquist(f):
            f = f.copy()
            f[:, :, NX // 2, :] = 0.0
            f[:, :, :, NZ // 2] = 0.0
            return f

        c.dist = lambda a, b: full(no_nyquist(a), no_nyquist(b))
        err_os, _ = c.run(d2.advance_resident, d2.state, nonlinearity)
        print("  device run, Orr-Sommerfeld measure without the Nyquist modes", err_os, flush=True)
    else:
        err_os, _ = c.run(d2.advance_resident, d2.state, nonlinearity)
        print("  device run, Orr-Sommerfeld measure", err_os, flush=True)
    assert err_os < 1e-7, err_os


@pytest.mark.parametrize("stepping", SCHEMES)
@pytest.mark.parametrize("base", [False, True], ids=["fluc", "base"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_gpu_parabola_is_steady(gpu, stepping, base, constraint):
    Nx, Ny, Nz = 8, 33, 8
    nu = 1.0 / 400
    g = R.Grid(Nx, Ny, Nz, 2 * np.pi, 2 * np.pi)
    fl = cfa.DNSFlags(timestepping=stepping, constraint=constraint, dealiasing="DealiasXZ",
                      nonlinearity="Rotational", baseflow="LaminarBase" if base else "ZeroBase", dt=0.1, nu=nu,
                      dPdx=-2.0 * nu, Ubulk=2.0 / 3.0)
    parab = np.zeros(Ny)
    parab[0], parab[2] = 0.5, -0.5
    zero = np.zeros(Ny)
    print("parabola", stepping, "base" if base else "fluc", constraint, flush=True)
    d = _gpu_dns(g, fl, g.zeros(3), g.zeros(1), (parab if base else zero, zero))
    u = g.zeros(3)
    if not base:
        u[0, :, 0, 0] += parab
    ut = u.copy()
    d.set_state([cfa.FlowField(Nx, Ny, Nz, 3, g.Lx, g.Lz, data=u), cfa.FlowField(Nx, Ny, Nz, 1, g.Lx, g.Lz)])
    err = 0.0
    for _ in range(10):
        err += np.sqrt(R.l2dist2(g, d.state()[0], ut))
        d.advance_resident(10)
    print("  FinalError", err, "dPdx", d.dPdx(), flush=True)
    assert err < 1e-13, err
    assert abs(d.dPdx() - (-2.0 * nu)) < 1e-13


@pytest.mark.parametrize("stepping", [s for s in SCHEMES if s != "SBDF1"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_gpu_zero_stays_zero(gpu, stepping, constraint):
    Nx, Ny, Nz = 8, 33, 8
    g = R.Grid(Nx, Ny, Nz, 2 * np.pi, 2 * np.pi)
    fl = cfa.DNSFlags(baseflow="LaminarBase", timestepping=stepping, constraint=constraint, dealiasing="DealiasXZ",
                      nonlinearity="Rotational", nu=1.0 / 400, dt=0.1)
    print("zero", stepping, constraint, flush=True)
    d = _gpu_dns(g, fl, g.zeros(3), g.zeros(1))
    err = 0.0
    for k in range(10):
        u = d.state()[0]
        err += np.sqrt(R.l2norm2(g, u))
        assert d.CFL() <= 2.0 and np.sqrt(R.l2norm2(g, u)) <= 1.0
        if k < 9:
            d.advance_resident(10)
    print("  FinalError", err, flush=True)
    assert err < ZERO_MAXERR[stepping], err
    assert d.dPdx() == 0.0
