"""The Orr-Sommerfeld eigenmode case of channelflow's dnsOrrsommTest
(tests/dnsOrrsommTest.cpp:39-291), shared by the oracle and GPU tests.

Set-up (dnsOrrsommTest.cpp:47-81): 4 x 65 x 4 grid, Lx = 2 pi, Lz = pi,
Re = 7500, dt = 0.02, SBDF3 started by CNRK2, NoDealiasing, constraint
BulkVelocity (Ubulk = 2/3) or PressureGradient (dPdx = -2 nu). The
eigenfunction (u, v, w and p at kx = 1, kz = 0, values at the Chebyshev
points, tests/data/os_{u,p}eig10_65.asc) is scaled so that its L2 norm is
1e-4 of the parabola's (:190-204) and added with its complex conjugate at
kx = -1. "base": the parabola is the base flow (ArbitraryBase) and u is the
eigenmode; "fluc": no base flow and u is the parabola plus the eigenmode
(:206-233). The linear solution at time t is the mode times exp(-i omega t).

The test's measure (:249-262) is the sum, over t = 0, 1, ..., 13 (every 50
steps), of L2Dist(u_n, u_l) + L2Dist(p_n, p_l), where p_n is the DNS's
physical pressure: p = q - |u + U|^2 / 2 for the rotational form and p = q
otherwise (DNS::uq2p, dns.cpp:372-406). The initial q is p plus |u + U|^2 / 2
for the rotational form, as DNS::up2q's own comment states (dns.cpp:408-441;
its code adds the half square of u + U's Chebyshev coefficients, with U in
place of W, to a zero q, which no longer matches that comment).
"""
import json
import os

import numpy as np

import cfref as R

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NZ = 4, 65, 4
LX, LZ = 2 * np.pi, np.pi
RE = 7500.0
NU = 1.0 / RE
DT = 0.02
NPLOT = 50          # steps between samples (dnsOrrsommTest.cpp:64)
T1 = 13.0
SCALE = 1e-4


def _spectral(vals):
    """Complex values at the Chebyshev points -> coefficients (ChebyTransform)."""
    re = np.ascontiguousarray(np.real(vals), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(vals), dtype=np.float64).copy()
    R.lib().cpuref_cheb_spectral(len(re), R.ptr(re))
    R.lib().cpuref_cheb_spectral(len(im), R.ptr(im))
    return re + 1j * im


class OSCase:
    def __init__(self, base=True):
        d = json.load(open(os.path.join(HERE, "golden", "os_eig10_65.json")))
        assert d["u"]["Ny"] == NY and d["u"]["kx"] == 1 and d["u"]["kz"] == 0 and d["u"]["state"] == "P"
        self.omega = complex(*d["omega"])
        self.lam = -1j * self.omega
        self.g = R.Grid(NX, NY, NZ, LX, LZ)
        self.base = base
        ue = [_spectral(np.array([complex(*v) for v in comp])) for comp in d["u"]["values"]]
        pe = [_spectral(np.array([complex(*v) for v in comp])) for comp in d["p"]["values"]]
        y = np.cos(np.pi * np.arange(NY) / (NY - 1))
        parab = (1.0 - y * y).astype(np.float64)
        R.lib().cpuref_cheb_spectral(NY, R.ptr(parab))
        self.parabola = parab
        # scale: perturbation norm = SCALE x the parabola field's norm (:190-204)
        ut = self.g.zeros(3)
        ut[0, :, 0, 0] += parab
        n_par = np.sqrt(R.l2norm2(self.g, ut, padded=False))
        ut = self.g.zeros(3)
        self._add_mode(ut, ue)
        n_pert = np.sqrt(R.l2norm2(self.g, ut, padded=False))
        c = SCALE * n_par / n_pert
        self.ue = [c * v for v in ue]
        self.pe = [c * v for v in pe]
        self.Ubase = parab.copy() if base else np.zeros(NY)
        self.Wbase = np.zeros(NY)

    @staticmethod
    def _add_mode(f, mode, factor=1.0):
        """f += mode at (kx=1, kz=0) and its conjugate at kx = -1 (FlowField += BasisFunc)."""
        for c, prof in enumerate(mode):
            f[c, :, 1, 0] += factor * prof
            f[c, :, NX - 1, 0] += np.conj(factor * prof)

    def flags(self, nonlinearity, constraint):
        return dict(baseflow="ArbitraryBase" if self.base else "ZeroBase", dealiasing="NoDealiasing",
                    nonlinearity=nonlinearity, constraint=constraint, timestepping="SBDF3",
                    initstepping="CNRK2", ulowerwall=0.0, uupperwall=0.0, dt=DT, nu=NU, dPdx=-2.0 * NU,
                    Ubulk=2.0 / 3.0)

    def exact(self, t):
        """(u_l, p_l) at time t: the linear solution (:264-276)."""
        f = np.exp(self.lam * t)
        u = self.g.zeros(3)
        self._add_mode(u, self.ue, f)
        if not self.base:
            u[0, :, 0, 0] += self.parabola
        p = self.g.zeros(1)
        self._add_mode(p, self.pe, f)
        return u, p

    def _half_square(self, u):
        """|u + U|^2 / 2 as a spectral 1-component field."""
        phys = R.field_physical(self.g, u)
        y = np.cos(np.pi * np.arange(NY) / (NY - 1))
        U = np.polynomial.chebyshev.chebval(y, self.Ubase)
        W = np.polynomial.chebyshev.chebval(y, self.Wbase)
        hs = 0.5 * ((phys[0] + U[:, None, None]) ** 2 + phys[1] ** 2 + (phys[2] + W[:, None, None]) ** 2)
        return R.field_spectral(self.g, hs[None])

    def up2q(self, u, p, nonlinearity):
        return p + self._half_square(u) if nonlinearity == "Rotational" else p.copy()

    def uq2p(self, u, q, nonlinearity):
        return q - self._half_square(u) if nonlinearity == "Rotational" else q.copy()

    def initial(self, nonlinearity):
        u0, p0 = self.exact(0.0)
        return u0, self.up2q(u0, p0, nonlinearity)

    def dist(self, a, b):
        return np.sqrt(R.l2dist2(self.g, a, b, padded=False))

    def run(self, advance, state, nonlinearity, samples=None):
        """The test loop: advance(n) steps the integration under test by n
        steps and state() returns its (u, q). Returns (err, per-sample list)."""
        err = 0.0
        per = []
        nsamp = int(round(T1)) + 1 if samples is None else samples
        for k in range(nsamp):
            t = float(k) * NPLOT * DT
            u, q = state()
            ul, pl = self.exact(t)
            pn = self.uq2p(u, q, nonlinearity) if k > 0 else self.exact(0.0)[1]
            ud, pd = self.dist(u, ul), self.dist(pn, pl)
            per.append((t, ud, pd))
            err += ud + pd
            if k + 1 < nsamp:
                advance(NPLOT)
        return err, per
