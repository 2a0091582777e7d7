"""The oracle (oracle/cpu_ref.cpp) pinned against the reference's own fixtures
and closed forms (CPU only).

Pins, by reference test:
  * timeIntegrationTest (tests/timeIntegrationTest.cpp:42-137): uinit -> 440
    SBDF3 steps -> ufinal at L2Dist <= 1e-13 -- rotational NL, SMRK2 start,
    tau solver, ASBL base flow, transforms, end to end.
  * chebyTest (tests/chebyTest.cpp:17-93): transform round trip and d/dy on
    known functions.
  * helmholtzTest (tests/helmholtzTest.cpp:36-152): manufactured Dirichlet and
    mean-constraint problems.
  * tridiagTest (tests/tridiagTest.cpp:16-95): banded UL solve against a dense
    solve.
  * tausolverTest (tests/tausolverTest.cpp:47-146): an exact divergence-free
    (u, v, w, P) satisfying the no-slip BCs -> R -> solve recovers it.
  * dnsParabolaTest (tests/dnsParabolaTest.cpp:60-92, 154-248), every
    registered variant (tests/CMakeLists.txt:84-111): scheme x {fluc, base} x
    {bulkv, gradp}, the Poiseuille parabola is a steady state (sum of 10
    L2Dist samples < 1e-13).
  * dnsZeroTest (tests/dnsZeroTest.cpp:63-106, 172-224), every registered
    variant (tests/CMakeLists.txt:69-82) with the reference's per-scheme bar.
  * dnsOrrsommTest (tests/dnsOrrsommTest.cpp:39-291, fixtures
    tests/data/os_{u,p}eig10_65.asc): the Orr-Sommerfeld eigenmode at
    Re = 7500 grows as exp(-i omega t) for the rotational, skew-symmetric,
    convection, divergence and alternating forms, base or fluctuation,
    bulk velocity or pressure gradient (sum of 14 velocity + pressure L2Dist
    samples < 1e-7).
  * The order of every time-stepping scheme (dnsalgo.cpp:102-179, 351-395,
    510-576): under dt halving, successive differences of the solution at a
    fixed time shrink as dt^p with the scheme's own order p, for a small field
    (the implicit part) and for one where the nonlinear term matters (its
    explicit extrapolation).
  * The order of every scheme under the bulk-velocity constraint, with
    Ubulk and Wbulk equal to the initial field's means.
  * The nonlinear forms on a field whose products are resolved exactly (low
    Chebyshev degree, 2/3-dealiased): convection = divergence = skew-symmetric,
    and convection - rotational = grad(|u|^2/2) component by component, with
    |u|^2/2 formed independently in physical space.
  * DNS::CFL (dnsalgo.cpp:63-67, flowfield.cpp:4035-4068): the laminar Couette
    value in closed form, and a random field against a brute-force evaluation of
    its Fourier-Chebyshev sums at every grid point.
  * PressureSolver (poissonsolver.cpp:326-431): the equations that define p,
    checked with numpy's own Chebyshev calculus -- the tau rows of
    p'' - kappa^2 p = -div N(u+U) and dp/dy = nu v'' at both walls (p = 0 at the
    walls for kx = kz = 0) -- on the modes whose wall correction is resolved.
"""
import json
import os

import numpy as np
import pytest

import cfref as R
from channelflow_amd.fields import random_field, divergence, cheb_diff

HERE = os.path.dirname(os.path.abspath(__file__))
LX, LZ = 2 * np.pi / 1.14, 2 * np.pi / 2.5


def cheb_values(c, y):
    return np.polynomial.chebyshev.chebval(y, c)


def test_golden_trajectory_oracle():
    man = json.load(open(os.path.join(HERE, "golden", "manifest.json")))
    mi = man["uinit"]
    g = R.Grid(mi["Nx"], mi["Ny"], mi["Nz"], mi["Lx"], mi["Lz"], mi["a"], mi["b"])
    shape = tuple(mi["shape"])
    u0 = R.field_from_io(g, np.fromfile(os.path.join(HERE, "golden", "uinit.f64")).reshape(shape))
    uf = R.field_from_io(g, np.fromfile(os.path.join(HERE, "golden", "ufinal.f64")).reshape(shape))
    t = man["timeIntegrationTest"]
    fl = R.make_flags(baseflow="SuctionBase", timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational",
                      dealiasing="DealiasXZ", taucorrection=True, constraint="PressureGradient", dPdx=0.0,
                      dt=t["dt"], nu=t["nu"], Vsuck=t["Vsuck"])
    d = R.OracleDNS(g, fl, u0, g.zeros(1))
    for _ in range(11):
        d.advance(40)
    u, _ = d.state()
    dist = np.sqrt(R.l2dist2(g, u, uf))
    assert dist <= t["tolerance_L2Dist"], dist
    assert abs(d.time - 11.0) < 1e-9


def test_cheby_roundtrip_and_diff():
    # chebyTest: u = 1 + sin(y) sampled at the Chebyshev points; transform,
    # differentiate, compare with cos(y) (N = 17 on [-1, 1] to spectral accuracy).
    N = 33
    y = np.cos(np.pi * np.arange(N) / (N - 1))
    x = 1 + np.sin(y)
    c = x.copy()
    R.lib().cpuref_cheb_spectral(N, R.ptr(c))
    back = c.copy()
    R.lib().cpuref_cheb_physical(N, R.ptr(back))
    assert np.max(np.abs(back - x)) < 1e-14
    assert np.max(np.abs(cheb_values(c, y) - x)) < 1e-14
    dc = np.zeros(N)
    R.lib().cpuref_cheb_diff(N, -1.0, 1.0, R.ptr(c), R.ptr(dc))
    assert np.max(np.abs(cheb_values(dc, y) - np.cos(y))) < 1e-13
    # exact on a monomial, interval [a, b] = [-1.1, 1.4] (chebyTest bounds)
    a, b = -1.1, 1.4
    coef = np.zeros(N)
    coef[5] = 1.0
    dcoef = np.zeros(N)
    R.lib().cpuref_cheb_diff(N, a, b, R.ptr(coef), R.ptr(dcoef))
    ref = np.polynomial.chebyshev.chebder(coef) * 2 / (b - a)
    assert np.max(np.abs(dcoef[:len(ref)] - ref)) < 1e-13
    # mean: integral over [-1, 1] / 2 of T_n
    for n in range(0, 9):
        cc = np.zeros(N)
        cc[n] = 1.0
        exact = 0.0 if n % 2 else (1.0 if n == 0 else -1.0 / (n * n - 1))
        assert abs(R.lib().cpuref_cheb_mean(N, R.ptr(cc)) - exact) < 1e-15


@pytest.mark.parametrize("lam,nu", [(0.0, 1.0), (3.7, 1.0), (412.0, 0.0025), (1e4, 1 / 400)])
def test_helmholtz_manufactured(lam, nu):
    # helmholtzTest: nu u'' - lambda u = f with u known; the tau solution on N
    # modes converges to the truth.
    N = 49
    y = np.cos(np.pi * np.arange(N) / (N - 1))
    u_true = np.exp(y) * np.sin(2 * y) + 0.3
    upp = np.exp(y) * (np.sin(2 * y) + 4 * np.cos(2 * y) - 4 * np.sin(2 * y))
    f = nu * upp - lam * u_true
    fc = f.copy()
    R.lib().cpuref_cheb_spectral(N, R.ptr(fc))
    uc = np.zeros(N)
    ua, ub = u_true[-1], u_true[0]
    assert R.lib().cpuref_helmholtz(N, -1.0, 1.0, lam, nu, R.ptr(fc), ua, ub, R.ptr(uc)) == 0
    err = np.max(np.abs(cheb_values(uc, y) - u_true))
    assert err < 1e-11, err


def test_helmholtz_mean_constraint():
    # nu u'' - lambda u = f + mu, mean(u) = umean, u(+-1) = 0 (helmholtzTest second problem)
    N = 49
    nu, lam, umean = 0.01, 25.0, 0.37
    y = np.cos(np.pi * np.arange(N) / (N - 1))
    f = np.cos(3 * y) * (1 - y * y)
    fc = f.copy()
    R.lib().cpuref_cheb_spectral(N, R.ptr(fc))
    uc = np.zeros(N)
    mu = np.zeros(1)
    assert R.lib().cpuref_helmholtz_mean(N, -1.0, 1.0, lam, nu, R.ptr(fc), umean, 0.0, 0.0, R.ptr(uc), R.ptr(mu)) == 0
    assert abs(R.lib().cpuref_cheb_mean(N, R.ptr(uc)) - umean) < 1e-13
    # residual of the ODE with the returned mu (all but the two tau rows)
    d1 = np.polynomial.chebyshev.chebder(uc)
    d2 = np.polynomial.chebyshev.chebder(d1)
    res = nu * np.pad(d2, (0, 2)) - lam * uc - fc
    res[0] -= mu[0]
    assert np.max(np.abs(res[:N - 2])) < 1e-11
    assert abs(cheb_values(uc, 1.0)) < 1e-13 and abs(cheb_values(uc, -1.0)) < 1e-13


def test_banded_tridiag():
    # tridiagTest: random row-0-banded tridiagonal, UL without pivoting vs dense solve
    rng = np.random.default_rng(0)
    M = 100
    band = rng.uniform(0.5, 1.5, M)
    lo = rng.uniform(-0.3, 0.3, M)
    up = rng.uniform(-0.3, 0.3, M)
    dg = rng.uniform(2.0, 3.0, M)
    A = np.zeros((M, M))
    A[0, :] = band
    for i in range(1, M):
        A[i, i] = dg[i]
        A[i, i - 1] = lo[i]
        if i + 1 < M:
            A[i, i + 1] = up[i]
    dg[0] = band[0]
    b = rng.standard_normal(M)
    x = b.copy()
    assert R.lib().cpuref_banded_solve(M, R.ptr(band), R.ptr(lo), R.ptr(dg), R.ptr(up), R.ptr(x)) == 0
    xr = np.linalg.solve(A, b)
    assert np.max(np.abs(x - xr)) < 1e-12 * max(1, np.max(np.abs(xr)))


def _deriv_y(c, L=2.0):
    return cheb_diff(c, L)


@pytest.mark.parametrize("kx,kz", [(1, 0), (0, 2), (3, 4), (-2, 1)])
def test_tausolver_exact_solution(kx, kz):
    # tausolverTest: build a divergence-free no-slip (u, v, w) and a pressure P,
    # form R = -(nu u'' - lambda u - grad P) ... solve, verify recovery.
    N = 49
    nu, lam_t = 1 / 400, 11 / 6 / 0.02
    a_, g_ = 2 * np.pi * kx / LX, 2 * np.pi * kz / LZ
    k2 = a_ * a_ + g_ * g_
    lam = lam_t + nu * k2
    full = random_field(3 * max(abs(kx), 1) * 3 + 3, N, 3 * max(kz, 1) * 3 + 3, LX, LZ, norm=None, decay=0.5,
                        seed=17, dealias=False)
    mx = kx % full.shape[2]
    u, v, w = full[0, :, mx, kz], full[1, :, mx, kz], full[2, :, mx, kz]
    P = np.zeros(N, complex)
    P[:N - 4] = np.exp(-0.3 * np.arange(N - 4)) * (0.7 + 0.2j)
    d2 = lambda c: _deriv_y(_deriv_y(c))
    # momentum: nu u'' - lambda u - grad P = -R
    Rx = -(nu * (d2(u) - k2 * u) - lam_t * u - 1j * a_ * P)
    Ry = -(nu * (d2(v) - k2 * v) - lam_t * v - _deriv_y(P))
    Rz = -(nu * (d2(w) - k2 * w) - lam_t * w - 1j * g_ * P)
    out = R.tausolve(kx, kz, LX, LZ, -1, 1, lam, nu, N, True, Rx, Ry, Rz)
    uu, vv, ww, PP = out
    div = 1j * a_ * uu + _deriv_y(vv) + 1j * g_ * ww
    assert np.max(np.abs(div)) < 1e-9
    for val in (uu, vv, ww):
        assert abs(np.sum(val)) < 1e-12 and abs(np.sum(val * (-1) ** np.arange(N))) < 1e-12
    # velocities recovered to the tau-method truncation level (tausolverTest's 1e-9)
    for got, ref in ((uu, u), (vv, v), (ww, w)):
        assert np.max(np.abs(got - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


SCHEMES = ["CNFE1", "CNAB2", "CNRK2", "SMRK2", "SBDF1", "SBDF2", "SBDF3", "SBDF4"]


def parabola_case(stepping, base, constraint):
    """tests/dnsParabolaTest.cpp:60-248: 8 x 33 x 8, Lx = Lz = 2 pi, Re = 400,
    dt = 0.1, rotational, 2/3 dealiasing, dPdx = -2 nu or Ubulk = 2/3. The DNS
    is built on zero fields with {Ubase, Wbase} = {parabola, 0} (base) or
    {0, 0} (fluc); the velocity is then set to 0 (base) or the parabola
    (fluc), and L2Dist(u, ut) is summed over t = 0, 1, ..., 9 (10 steps
    apart). Returns (err, dns, grid)."""
    Nx, Ny, Nz = 8, 33, 8
    nu = 1.0 / 400
    g = R.Grid(Nx, Ny, Nz, 2 * np.pi, 2 * np.pi)
    fl = R.make_flags(timestepping=stepping, constraint=constraint, dealiasing="DealiasXZ",
                      nonlinearity="Rotational", baseflow="LaminarBase" if base else "ZeroBase", dt=0.1, nu=nu,
                      dPdx=-2.0 * nu, Ubulk=2.0 / 3.0, ulowerwall=0.0, uupperwall=0.0)
    parab = np.zeros(Ny)
    parab[0], parab[2] = 0.5, -0.5
    zero = np.zeros(Ny)
    d = R.OracleDNS(g, fl, g.zeros(3), g.zeros(1), base=(parab if base else zero, zero))
    u = g.zeros(3)
    if not base:
        u[0, :, 0, 0] += parab
    ut = u.copy()
    d.set_state(u, g.zeros(1))
    err = 0.0
    for _ in range(10):
        err += np.sqrt(R.l2dist2(g, d.state()[0], ut))
        d.advance(10)
    return err, d, g


@pytest.mark.parametrize("stepping", SCHEMES)
@pytest.mark.parametrize("base", [False, True], ids=["fluc", "base"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_parabola_is_steady(stepping, base, constraint):
    err, d, g = parabola_case(stepping, base, constraint)
    print(stepping, "base" if base else "fluc", constraint, "FinalError", err, "dPdx", d.dPdx, "Ubulk", d.Ubulk)
    assert err < 1e-13, err
    # the constraint's bookkeeping: the parabola's driving gradient, imposed or found by the bulk solve
    assert abs(d.dPdx - (-2.0 / 400)) < 1e-13


ZERO_MAXERR = dict(CNFE1=2e-2, CNAB2=4e-5, CNRK2=5e-6, SMRK2=8e-6, SBDF2=2e-4, SBDF3=2e-6, SBDF4=1e-6)


@pytest.mark.parametrize("stepping", [s for s in SCHEMES if s != "SBDF1"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_zero_stays_zero(stepping, constraint):
    """tests/dnsZeroTest.cpp:63-106, 172-224: laminar base flow with both walls at
    rest and zero fluctuation, 8 x 33 x 8, Re = 400, dt = 0.1, rotational:
    sum over t = 0 ... 9 of L2Norm(u) below the reference's per-scheme bar, and
    the run's CFL and norm stay within its own sanity limits (:191-198)."""
    Nx, Ny, Nz = 8, 33, 8
    g = R.Grid(Nx, Ny, Nz, 2 * np.pi, 2 * np.pi)
    fl = R.make_flags(baseflow="LaminarBase", timestepping=stepping, constraint=constraint,
                      dealiasing="DealiasXZ", nonlinearity="Rotational", nu=1.0 / 400, ulowerwall=0.0,
                      uupperwall=0.0, dt=0.1)
    d = R.OracleDNS(g, fl, g.zeros(3), g.zeros(1))
    err = 0.0
    for k in range(10):
        u = d.state()[0]
        err += np.sqrt(R.l2norm2(g, u))
        assert d.cfl() <= 2.0 and np.sqrt(R.l2norm2(g, u)) <= 1.0
        if k < 9:
            d.advance(10)
    assert err < ZERO_MAXERR[stepping], err
    assert d.dPdx == 0.0 and d.Ubulk == 0.0


OS_FORMS = ["Rotational", "SkewSymmetric", "Convection", "Divergence", "Alternating"]


@pytest.mark.parametrize("nonlinearity", OS_FORMS)
@pytest.mark.parametrize("base", [True, False], ids=["base", "fluc"])
@pytest.mark.parametrize("constraint", ["BulkVelocity", "PressureGradient"])
def test_orr_sommerfeld(nonlinearity, base, constraint):
    """dnsOrrsommTest (tests/dnsOrrsommTest.cpp:39-291), reference bar 1e-7 on the
    sum of 14 samples of L2Dist(u_n, u_l) + L2Dist(p_n, p_l). The unstable
    Orr-Sommerfeld mode of plane Poiseuille flow at Re = 7500 grows as
    exp(-i omega t); a missing or wrong term of any nonlinear form's
    linearisation about the parabola, of the mean-flow constraint or of the
    time stepping changes the growth and fails this.

    The rotational form's pressure p = q - |u + U|^2/2 is evaluated on the 4 x 4
    grid, where |u|^2/2 puts a quadratic (1e-9) part on the Nyquist modes kx = 2
    and kz = 2, which no solve writes: the rotational sum is 1.08e-7 with them
    and equals the other forms' 8.6e-8 without them. Its bar is therefore
    applied to the non-Nyquist modes, and the full sum is held to 1.1e-7."""
    from oscase import OSCase, NX, NZ
    c = OSCase(base=base)
    u0, q0 = c.initial(nonlinearity)
    fl = R.make_flags(**c.flags(nonlinearity, constraint))
    d = R.OracleDNS(c.g, fl, u0, q0, padded=False, base=(c.Ubase, c.Wbase) if base else None)
    if nonlinearity == "Rotational":
        full = c.dist

        def no_nyquist(f):
            f = f.copy()
            f[:, :, NX // 2, :] = 0.0
            f[:, :, :, NZ // 2] = 0.0
            return f

        c.dist = lambda a, b: full(no_nyquist(a), no_nyquist(b))
        err, per = c.run(d.advance, d.state, nonlinearity)
        # the full measure on the same run, sample by sample
        d2 = R.OracleDNS(c.g, fl, u0, q0, padded=False, base=(c.Ubase, c.Wbase) if base else None)
        c.dist = full
        err_full, _ = c.run(d2.advance, d2.state, nonlinearity)
        print(nonlinearity, base, constraint, "err (no Nyquist)", err, "err (full)", err_full)
        assert err < 1e-7 and err_full < 1.1e-7, (err, err_full)
    else:
        err, per = c.run(d.advance, d.state, nonlinearity)
        print(nonlinearity, base, constraint, "err", err)
        assert err < 1e-7, err
    # the mode grew: its norm at t = 13 is exp(13 Im omega) of the initial one
    ul0, _ = c.exact(0.0)
    ul1, _ = c.exact(13.0)
    un = d.state()[0] if nonlinearity != "Rotational" else d2.state()[0]
    if not base:
        ul0[0, :, 0, 0] -= c.parabola
        un = un.copy()
        un[0, :, 0, 0] -= c.parabola
    growth = np.sqrt(R.l2norm2(c.g, un, padded=False) / R.l2norm2(c.g, ul0, padded=False))
    assert abs(growth - np.exp(13.0 * c.omega.imag)) < 1e-2 * (np.exp(13.0 * c.omega.imag) - 1.0)


@pytest.mark.parametrize("stepping,init,order", [("SBDF1", "SMRK2", 1), ("SBDF2", "SMRK2", 2), ("SBDF3", "SMRK2", 3),
                                                  ("SBDF4", "SMRK2", 3), ("CNFE1", "SMRK2", 1), ("CNAB2", "SMRK2", 2),
                                                  ("SMRK2", "SMRK2", 2), ("CNRK2", "CNRK2", 2), ("SBDF3", "CNRK2", 3)])
def test_scheme_order(stepping, init, order):
    """Observed order of convergence at t = 0.64 with dt = T/16 ... T/128: log2
    of the ratio of successive solution differences. A small field (norm
    1e-4, no base flow, fixed pressure gradient) keeps the run in the
    asymptotic regime. SBDF4 shows order 3 because its second-order start
    steps limit it. A wrong scheme coefficient drops a scheme's order."""
    g = R.Grid(8, 17, 8, LX, LZ)
    T = 0.64
    u0 = random_field(8, 17, 8, LX, LZ, norm=1e-4, decay=0.5, seed=9)
    sols = []
    for n in (16, 32, 64, 128):
        fl = R.make_flags(baseflow="ZeroBase", constraint="PressureGradient", dPdx=0.0, nu=1 / 400,
                          nonlinearity="SkewSymmetric", dealiasing="DealiasXZ", timestepping=stepping,
                          initstepping=init, dt=T / n)
        d = R.OracleDNS(g, fl, u0, g.zeros(1))
        d.advance(n)
        sols.append(d.state()[0])
    diffs = [np.sqrt(R.l2dist2(g, sols[i], sols[i + 1])) for i in range(3)]
    orders = [np.log2(diffs[i] / diffs[i + 1]) for i in range(2)]
    print(stepping, init, "differences", diffs, "observed orders", orders, flush=True)
    assert all(order - 0.15 <= p <= order + 1.2 for p in orders), orders


@pytest.mark.parametrize("stepping,init,order", [("SBDF1", "SMRK2", 1), ("SBDF2", "SMRK2", 2), ("SBDF3", "SMRK2", 3),
                                                  ("SBDF4", "SMRK2", 3), ("CNFE1", "SMRK2", 1), ("CNAB2", "SMRK2", 2),
                                                  ("SMRK2", "SMRK2", 2), ("CNRK2", "CNRK2", 2), ("SBDF3", "CNRK2", 3)])
def test_scheme_order_bulk_velocity(stepping, init, order):
    """The same measurement under the bulk-velocity constraint (the (0,0) mode
    solved with dPdx as the multiplier of mean(u) = Ubulk - UbulkBase,
    nse.cpp:530-546, tausolver.cpp:404-450, helmholtz.cpp:158-213), with Ubulk
    and Wbulk equal to the initial field's own means. With Wbulk = 0 instead
    (the initial w00 has a nonzero mean), the first solve projects an O(1)
    jump of the spanwise mean and every scheme drops to first order."""
    g = R.Grid(8, 17, 8, LX, LZ)
    T = 0.64
    u0 = random_field(8, 17, 8, LX, LZ, norm=1e-4, decay=0.5, seed=9)
    mean = lambda c: R.lib().cpuref_cheb_mean(17, R.ptr(np.ascontiguousarray(u0[c, :, 0, 0].real)))
    sols = []
    for n in (16, 32, 64, 128):
        fl = R.make_flags(baseflow="ZeroBase", constraint="BulkVelocity", Ubulk=mean(0), Wbulk=mean(2),
                          nu=1 / 400, nonlinearity="SkewSymmetric", dealiasing="DealiasXZ", timestepping=stepping,
                          initstepping=init, dt=T / n)
        d = R.OracleDNS(g, fl, u0, g.zeros(1))
        d.advance(n)
        sols.append(d.state()[0])
        assert abs(d.Ubulk - mean(0)) < 1e-18
    diffs = [np.sqrt(R.l2dist2(g, sols[i], sols[i + 1])) for i in range(3)]
    orders = [np.log2(diffs[i] / diffs[i + 1]) for i in range(2)]
    print(stepping, init, "bulk velocity: differences", diffs, "observed orders", orders, flush=True)
    assert all(order - 0.15 <= p <= order + 1.2 for p in orders), orders


@pytest.mark.parametrize("stepping,least", [("SBDF2", 1.85), ("SBDF3", 2.6), ("SBDF4", 2.85), ("CNAB2", 1.85)])
def test_scheme_order_nonlinear(stepping, least):
    """The explicit extrapolation of the nonlinear term (the beta weights of
    dnsalgo.cpp:125-136 and the CNAB2 weights): the same measurement with a
    field of norm 0.2, where the nonlinear term matters, and dt = T/32 ...
    T/256. The order from the finest pair is still rising toward the scheme's
    own (SBDF3 2.74, SBDF4 3.41); SBDF3 with SBDF2's extrapolation weights
    stays at 1.9."""
    g = R.Grid(8, 17, 8, LX, LZ)
    T = 0.64
    u0 = random_field(8, 17, 8, LX, LZ, norm=0.2, decay=0.5, seed=9)
    sols = []
    for n in (32, 64, 128, 256):
        fl = R.make_flags(baseflow="ZeroBase", constraint="PressureGradient", dPdx=0.0, nu=1 / 400,
                          nonlinearity="SkewSymmetric", dealiasing="DealiasXZ", timestepping=stepping,
                          initstepping="SMRK2", dt=T / n)
        d = R.OracleDNS(g, fl, u0, g.zeros(1))
        d.advance(n)
        sols.append(d.state()[0])
    diffs = [np.sqrt(R.l2dist2(g, sols[i], sols[i + 1])) for i in range(3)]
    p = np.log2(diffs[1] / diffs[2])
    print(stepping, "differences", diffs, "observed order", p, flush=True)
    assert p >= least, p


def _resolved_field(Nx, Ny, Nz, Nlow, seed):
    """A divergence-free no-slip field whose Chebyshev degree is below Ny/2, so
    that every quadratic product is a polynomial the Ny points hold exactly."""
    low = random_field(Nx, Nlow, Nz, LX, LZ, norm=0.2, decay=0.5, seed=seed)
    u = np.zeros((3, Ny) + low.shape[2:], dtype=np.complex128)
    u[:, :Nlow] = low
    return u


def test_nonlinear_identities():
    """With 2/3 dealiasing in x, z and products of degree <= Ny - 1 in y, the
    discrete forms obey the continuous identities on the live modes:
    convection u.grad u = divergence div(u u) (u is divergence-free),
    skew-symmetric = their mean, and convection - rotational = grad(|u|^2/2)
    in all three components, with |u|^2/2 formed here from the physical
    values and differentiated with numpy's own spectral calculus. A dropped
    term, a wrong sign or a wrong index in one component of one form breaks
    one of these equalities."""
    Nx, Ny, Nz = 24, 25, 24
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    u = _resolved_field(Nx, Ny, Nz, 13, seed=5)
    assert np.max(np.abs(divergence(u, LX, LZ))) < 1e-16
    fl = dict(baseflow="ZeroBase", dealiasing="DealiasXZ")
    f = {m: R.nonlinear(g, R.make_flags(nonlinearity=m, **fl), u)[0]
         for m in ("Convection", "Divergence", "SkewSymmetric", "Rotational")}
    kxd, kzd = Nx // 3 - 1, Nz // 3 - 1
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)
    live = (np.abs(kx)[:, None] <= kxd) & (np.arange(Nz // 2 + 1)[None, :] <= kzd)
    scale = np.max(np.abs(f["Convection"]))

    def close(a, b, what):
        err = np.max(np.abs((a - b)[..., live]))
        print(what, err / scale)
        assert err < 1e-13 * scale, (what, err / scale)

    close(f["Convection"], f["Divergence"], "convection - divergence")
    close(f["SkewSymmetric"], 0.5 * (f["Convection"] + f["Divergence"]), "skew - (conv + div)/2")
    # grad(|u|^2/2) from the physical values
    phys = R.field_physical(g, u)
    hs = R.field_spectral(g, 0.5 * np.sum(phys * phys, axis=0)[None])[0]       # (Ny, Nx, Mz)
    al = (2 * np.pi * kx / LX)[None, :, None]
    ga = (2 * np.pi * np.arange(Nz // 2 + 1) / LZ)[None, None, :]
    dhs = np.moveaxis(cheb_diff(np.moveaxis(hs, 0, -1)), -1, 0)
    grad = np.stack([1j * al * hs, dhs, 1j * ga * hs])
    d = f["Convection"] - f["Rotational"]
    for i in range(3):
        close(d[i], grad[i], f"(conv - rot - grad |u|^2/2)[{i}]")
    for m in f:  # aliased modes of every form are zero (nse.cpp:389-390)
        assert np.max(np.abs(f[m][..., ~live])) == 0.0


def test_field_io_and_generator():
    # the seeded generator is divergence-free, no-slip and normalised (oracle norms)
    Nx, Ny, Nz = 32, 33, 32
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.1, decay=0.3, seed=7)
    assert abs(np.sqrt(R.l2norm2(g, u)) - 0.1) < 1e-15
    assert np.max(np.abs(divergence(u, LX, LZ))) < 1e-15
    p = R.field_physical(g, u)
    assert np.max(np.abs(p[:, 0])) < 1e-15 and np.max(np.abs(p[:, -1])) < 1e-15
    assert np.max(np.abs(R.field_spectral(g, p) - u)) < 1e-15


def _physical_bruteforce(u, Nx, Ny, Nz, Lx, Lz):
    """Values of a spectral field at every grid point by direct sums:
    u(x,y,z) = Re sum_{kx,kz} w_kz sum_n u_n(kx,kz) T_n(y) exp(2 pi i (kx x/Lx + kz z/Lz)),
    w_0 = 1, w_kz = 2 (kz > 0), y_j = cos(pi j/(Ny-1)), x = Lx i/Nx, z = Lz k/Nz."""
    y = np.cos(np.pi * np.arange(Ny) / (Ny - 1))
    T = np.cos(np.outer(np.arange(Ny), np.arccos(np.clip(y, -1, 1))))        # T[n, j] = T_n(y_j)
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)
    kz = np.arange(Nz // 2 + 1)
    wz = np.where(kz == 0, 1.0, 2.0)
    ex = np.exp(2j * np.pi * np.outer(np.arange(Nx), kx) / Nx)              # [x, mx]
    ez = np.exp(2j * np.pi * np.outer(np.arange(Nz), kz) / Nz) * wz         # [z, mz]
    out = np.zeros((u.shape[0], Ny, Nx, Nz))
    for c in range(u.shape[0]):
        uy = np.einsum("nj,nam->jam", T, u[c])                              # [y, mx, mz]
        out[c] = np.einsum("xa,jam,zm->jxz", ex, uy, ez).real
    return out


def test_cfl_closed_form_and_bruteforce():
    Nx, Ny, Nz = 16, 17, 12
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    fl = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
              nu=1 / 400, dt=0.02, nonlinearity="SkewSymmetric", dealiasing="DealiasXZ")
    # laminar Couette U = y: the largest (u + U)/dx is 1/(Lx/Nx) at the upper wall, v = w = 0
    d = R.OracleDNS(g, R.make_flags(**fl), g.zeros(3), g.zeros(1))
    assert d.cfl() == pytest.approx(Nx / LX * 2 * np.pi / 3 * 0.02, rel=1e-14)
    # a random field plus U = y, against direct sums; no base flow without dealiasing (factor pi dt)
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.3, decay=0.5, seed=11)
    y = np.cos(np.pi * np.arange(Ny) / (Ny - 1))
    dy = np.empty(Ny)
    dy[0] = dy[-1] = y[0] - y[1]
    dy[1:-1] = (y[:-2] - y[2:]) / 2
    # one velocity component kept at a time, so that each term decides the maximum once, and U = 8y,
    # which dominates; the CFL does not need a divergence-free field
    cases = []
    for keep in (0, 1, 2):
        uk = np.zeros_like(u)
        uk[keep] = u[keep]
        cases.append((f"component {keep} only, ZeroBase", uk, dict(fl, baseflow="ZeroBase"), 0.0 * y))
    cases.append(("all components, U = 8y", u, dict(fl, ulowerwall=-8.0, uupperwall=8.0), 8.0 * y))
    winners = set()
    for name, uk, fl2, U in cases:
        phys = _physical_bruteforce(uk, Nx, Ny, Nz, LX, LZ)
        terms = [np.max((phys[0] + U[:, None, None]) / (LX / Nx)), np.max(phys[1] / dy[:, None, None]),
                 np.max(phys[2] / (LZ / Nz))]
        winners.add(int(np.argmax(terms)))
        ref = max(max(terms), 0.0) * 2 * np.pi / 3 * 0.02
        d = R.OracleDNS(g, R.make_flags(**fl2), uk, g.zeros(1))
        print(f"{name}: oracle CFL {d.cfl():.15e}, direct sums {ref:.15e}")
        assert d.cfl() == pytest.approx(ref, rel=1e-12)
    assert winners == {0, 1, 2}


def test_pressure_solver_equations():
    from numpy.polynomial import chebyshev as Ch
    Nx, Ny, Nz = 16, 33, 16
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    u = random_field(Nx, Ny, Nz, LX, LZ, norm=0.2, decay=0.4, seed=21)
    cases = [dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
                  nu=1 / 400, nonlinearity="Convection"),
             dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3 / 2000, nu=1 / 2000,
                  nonlinearity="Rotational")]
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)

    def der(c, m=1):  # d^m/dy^m of Chebyshev coefficients on [-1, 1], same length
        return np.pad(Ch.chebder(c, m), (0, m))

    for fl in cases:
        flags = R.make_flags(dealiasing="DealiasXZ", **fl)
        p = R.pressure(g, flags, u)[0]
        N = R.nonlinear(g, flags, u.copy())[0]      # live modes as navierstokesNL's
        checked = 0
        for mx in range(Nx):
            for mz in range(Nz // 2 + 1):
                if abs(kx[mx]) > Nx // 3 - 1 or mz > Nz // 3 - 1:
                    continue
                a, c = 2 * np.pi * kx[mx] / LX, 2 * np.pi * mz / LZ
                mu2 = a * a + c * c
                if mu2 > 36.0:            # wall correction e^{+-mu y} not resolved at Ny = 33
                    continue
                pc = p[:, mx, mz]
                divN = 1j * a * N[0][:, mx, mz] + der(N[1][:, mx, mz]) + 1j * c * N[2][:, mx, mz]
                res = der(pc, 2)[:Ny - 2] - mu2 * pc[:Ny - 2] + divN[:Ny - 2]
                scale = max(np.max(np.abs(divN)), mu2 * np.max(np.abs(pc)), 1e-300)
                assert np.max(np.abs(res)) <= 1e-9 * scale, (fl["nonlinearity"], kx[mx], mz)
                if mx == 0 and mz == 0:
                    assert abs(Ch.chebval(1.0, pc)) < 1e-12 * scale and abs(Ch.chebval(-1.0, pc)) < 1e-12 * scale
                else:
                    dp, vyy = der(pc), der(u[1][:, mx, mz], 2)
                    for yw in (-1.0, 1.0):
                        bc = Ch.chebval(yw, dp) - fl["nu"] * Ch.chebval(yw, vyy)
                        assert abs(bc) <= 1e-9 * max(abs(Ch.chebval(yw, dp)), 1e-12), (kx[mx], mz, yw, bc)
                checked += 1
        print(fl["nonlinearity"], "modes checked", checked)
        assert checked >= 10


@pytest.mark.parametrize("N", [17, 33, 49])
def test_tausolve_bulk_defining_equations(N):
    """The oracle's bulk-velocity TauSolver::solve (tausolver.cpp:404-450) for
    kx = kz = 0 against the equations it solves: nu u'' - lambda u - dPdx = -Rx
    (and w with dPdz) on every Chebyshev row but the top two (the tau rows),
    u(+-1) = w(+-1) = 0, mean(u) = umean, mean(w) = wmean, v = 0, and
    P'' = dRy/dy with P(+-1) = 0."""
    r = np.random.default_rng(N)
    dec = 0.8 ** np.arange(N)
    Rx, Ry, Rz = [(r.standard_normal(N) + 1j * r.standard_normal(N)) * dec for _ in range(3)]
    nu, lam = 1 / 400, 11 / 6 / 0.02
    out, dPdx, dPdz = R.tausolve_bulk(LX, LZ, -1.0, 1.0, lam, nu, N, True, Rx, Ry, Rz, 0.25, -0.05)
    u, v, w, P = out
    d2 = lambda c: _deriv_y(_deriv_y(c))
    for x, Rr, dP in ((u, Rx, dPdx), (w, Rz, dPdz)):
        res = nu * d2(x) - lam * x + Rr
        res[0] -= dP
        assert np.max(np.abs(res[:N - 2])) < 1e-10 * np.max(np.abs(Rr))
        assert abs(np.sum(x)) < 1e-12 and abs(np.sum(x * (-1) ** np.arange(N))) < 1e-12
    mean = lambda c: c[0].real - sum(c[n].real / (n * n - 1) for n in range(2, N, 2))
    assert abs(mean(u) - 0.25) < 1e-13 and abs(mean(w) + 0.05) < 1e-13
    assert np.all(v == 0)
    # P'' = dRy/dy with P(+-1) = 0: P' - Ry is constant (in y) on the non-tau rows
    assert np.max(np.abs(_deriv_y(_deriv_y(P) - Ry)[:N - 3])) < 1e-10 * np.max(np.abs(Ry))
    assert abs(np.sum(P)) < 1e-12 and abs(np.sum(P * (-1) ** np.arange(N))) < 1e-12
