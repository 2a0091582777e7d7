"""Multiple-shooting batch client on the GPU (channelflow_amd.dsi).

The residual G_i = f^T(u_i) - u_{i+1} and the finite-difference Jacobian
products (G(x + eps dx) - G(x)) / eps computed from one Batch must equal the
same quantities built from separate single-field DNS runs bit for bit (a
member's result does not depend on the batch), and G must stay within the
20-step bar of the oracle (relative L2 <= 1e-10 per shot and component). The
Jacobian products are compared with the oracle's own finite differences at
1e-6 relative: the quotient divides integration differences of about 1e-13
by eps ~ 1e-6.
"""
import numpy as np
import pytest

import cfref as R
import channelflow_amd as cfa
from channelflow_amd.dsi import MultiShooting
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu

LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5
COUETTE = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
               nu=1 / 400, dt=0.02, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric",
               dealiasing="DealiasXZ", variabledt=False)
NX, NY, NZ = 16, 25, 16
T = 0.2  # 10 steps


def shots_and_dirs(nshot=3, ndir=2):
    shots = [random_field(NX, NY, NZ, LX_C, LZ_C, norm=0.1 + 0.05 * i, decay=0.4, seed=300 + i) for i in range(nshot)]
    dirs = [[random_field(NX, NY, NZ, LX_C, LZ_C, norm=0.1, decay=0.4, seed=400 + 10 * k + i) for i in range(nshot)]
            for k in range(ndir)]
    return shots, dirs


def single_ft(u, flags, nsteps):
    fu = cfa.FlowField(NX, NY, NZ, 3, LX_C, LZ_C, data=u)
    fq = cfa.FlowField(NX, NY, NZ, 1, LX_C, LZ_C)
    d = cfa.DNS([fu, fq], flags)
    d.advance_resident(nsteps)
    return d.state()[0]


def oracle_ft(u, flags, nsteps):
    g = R.Grid(NX, NY, NZ, LX_C, LZ_C)
    ref = R.OracleDNS(g, R.make_flags(**flags), u, g.zeros(1))
    ref.advance(nsteps)
    return ref.state()[0]


def rel(g, a, b):
    return max(np.sqrt(R.l2dist2(g, a[i:i + 1], b[i:i + 1]) / R.l2norm2(g, b[i:i + 1])) for i in range(3))


@pytest.mark.parametrize("centdiff", [False, True])
def test_residual_and_jacobians_match_single_runs(gpu, centdiff):
    shots, dirs = shots_and_dirs()
    fl = cfa.DNSFlags(**COUETTE)
    ms = MultiShooting(NX, NY, NZ, LX_C, LZ_C, fl, T)
    G, DG, eps = ms.eval_with_jacobians(shots, dirs, epsDx=1e-7, centdiff=centdiff)
    assert ms.nsteps == 10 and ms.fcount == 3 * (1 + 2 * (2 if centdiff else 1))
    n = len(shots)
    base = [single_ft(u, ms.flags, ms.nsteps) for u in shots]
    for i in range(n):
        assert np.array_equal(G[i], base[i] - shots[(i + 1) % n])

    def G_at(x):
        fx = [single_ft(u, ms.flags, ms.nsteps) for u in x]
        return [fx[i] - x[(i + 1) % n] for i in range(n)]

    for k, dx in enumerate(dirs):
        e = eps[k]
        if centdiff:
            Gp = G_at([shots[i] + (0.5 * e) * dx[i] for i in range(n)])
            Gm = G_at([shots[i] - (0.5 * e) * dx[i] for i in range(n)])
            expect = [(1 / e) * (Gp[i] - Gm[i]) for i in range(n)]
        else:
            Gp = G_at([shots[i] + e * dx[i] for i in range(n)])
            expect = [(1 / e) * (Gp[i] - G[i]) for i in range(n)]
        for i in range(n):
            assert np.array_equal(DG[k][i], expect[i])
    # eval alone gives the same G
    G2 = MultiShooting(NX, NY, NZ, LX_C, LZ_C, fl, T).eval(shots)
    for i in range(n):
        assert np.array_equal(G2[i], G[i])


def test_residual_and_jacobian_vs_oracle(gpu):
    shots, dirs = shots_and_dirs(nshot=2, ndir=1)
    fl = cfa.DNSFlags(**COUETTE)
    ms = MultiShooting(NX, NY, NZ, LX_C, LZ_C, fl, T)
    G, DG, eps = ms.eval_with_jacobians(shots, dirs, epsDx=1e-7)
    flags = dict(COUETTE, dt=ms.flags.dt)
    g = R.Grid(NX, NY, NZ, LX_C, LZ_C)
    n = len(shots)
    fo = [oracle_ft(u, flags, ms.nsteps) for u in shots]
    Go = [fo[i] - shots[(i + 1) % n] for i in range(n)]
    e = eps[0]
    xp = [shots[i] + e * dirs[0][i] for i in range(n)]
    fp = [oracle_ft(u, flags, ms.nsteps) for u in xp]
    DGo = [(1 / e) * ((fp[i] - xp[(i + 1) % n]) - Go[i]) for i in range(n)]
    for i in range(n):
        # f^T itself at the 20-step bar, G (a difference of O(1) fields) relative to f^T's size
        assert rel(g, G[i] + shots[(i + 1) % n], fo[i]) <= 1e-10
        ej = rel(g, DG[0][i], DGo[i])
        print(f"shot {i}: eps {e:.3e}, Jacobian product rel L2 vs oracle {ej:.2e}", flush=True)
        assert ej <= 1e-6


def test_residual_variable_dt_matches_single_and_oracle(gpu):
    """flags.variabledt: f^T as cfDSI::f computes it (cfdsi.cpp:733-767) -- a
    probe step whose CFL adjusts dt, a restart from u, then every dT the CFL at
    its start decides a reset_dt -- for every shot inside one Batch, equal bit
    for bit to a single DNS driven the same way, and within 1e-10 of the
    oracle driven the same way."""
    import copy
    shots, _ = shots_and_dirs(nshot=3, ndir=0)
    fl = cfa.DNSFlags(**dict(COUETTE, variabledt=True, dT=0.1, dtmax=0.05))
    ms = MultiShooting(NX, NY, NZ, LX_C, LZ_C, fl, 1.0)
    finals = ms.integrate(shots)
    g = R.Grid(NX, NY, NZ, LX_C, LZ_C)
    for i, u in enumerate(shots):
        dt = copy.deepcopy(ms.timestep)
        d = cfa.DNS([cfa.FlowField(NX, NY, NZ, 3, LX_C, LZ_C, data=u), cfa.FlowField(NX, NY, NZ, 1, LX_C, LZ_C)],
                    ms.flags)
        o = R.OracleDNS(g, R.make_flags(**dict(COUETTE, dt=ms.flags.dt)), u, g.zeros(1))
        d.advance_resident(1)
        o.advance(1)
        dt.adjust(d.CFL())
        d.reset_dt(dt.dt())
        o.reset_dt(dt.dt())
        d.set_state([cfa.FlowField(NX, NY, NZ, 3, LX_C, LZ_C, data=u)])
        o.set_state(u)
        seq = [dt.dt()]
        for _ in range(ms.timestep.N()):
            cfl, cflo = d.CFL(), o.cfl()
            assert abs(cfl - cflo) <= 1e-12 * cflo
            d.advance_resident(dt.n())
            o.advance(dt.n())
            if dt.adjust(cfl):
                d.reset_dt(dt.dt())
                o.reset_dt(dt.dt())
            seq.append(dt.dt())
        us = d.state()[0]
        print(f"shot {i}: dt sequence {seq}", flush=True)
        assert np.array_equal(finals[i], us)
        assert ms.dts[i] == dt.dt()
        assert rel(g, us, o.state()[0]) <= 1e-10
    assert any(x != COUETTE["dt"] for x in ms.dts)  # dt did adapt
