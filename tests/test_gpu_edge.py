"""Edge-tracking batch client on the GPU (channelflow_amd.edge).

Every attractor test of a batched call must reach the same verdict, after
the same number of time units, with the same final field bit for bit, as
CalcAttractor (programs/edgetracking.cpp:767-812) run one field at a time
with its own DNS; and a bisection round must pick the same uL and uH.
"""
import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd import edge as E
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu

LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5
NX, NY, NZ = 16, 17, 16
FLAGS = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
             nu=1 / 400, dt=0.02, dT=0.1, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric",
             dealiasing="DealiasXZ", variabledt=False)
AFLAGS = E.AttractorFlags(EcfH=4e-3, EcfL=2e-4)


def calc_attractor_single(u0, flags, aflags, nstep, tMin, tMax):
    """CalcAttractor with a DNS of its own (edgetracking.cpp:767-812, fixed dt)."""
    fu = cfa.FlowField(NX, NY, NZ, 3, LX_C, LZ_C, data=u0)
    fq = cfa.FlowField(NX, NY, NZ, 1, LX_C, LZ_C)
    cfa.PressureSolver(fu, flags=flags).solve(fq, fu)
    dns = cfa.DNS([fu, fq], flags)
    u = np.array(u0, dtype=np.complex128)
    result, t, tDecay = E.NO_ATTRACTOR, 0, 0
    while result == E.NO_ATTRACTOR and t < tMax:
        t += 1
        lastEcf = E.ecf(u)
        dns.advance_resident(nstep)
        u = dns.state()[0]
        if t > tMin:
            result, tDecay = E.classify(u, tDecay, lastEcf, aflags)
    return result, t, u


def members():
    return [random_field(NX, NY, NZ, LX_C, LZ_C, norm=n, decay=0.4, seed=700 + i)
            for i, n in enumerate([0.005, 0.3, 0.03, 0.12, 0.01])]


def test_attractors_match_single_runs(gpu):
    fl = cfa.DNSFlags(**FLAGS)
    tr = E.EdgeTracker(NX, NY, NZ, LX_C, LZ_C, fl, AFLAGS, tMin=1, tMax=12)
    assert tr.nstep == 5
    us = members()
    got = tr.attractors(us)
    kinds = set()
    for u0, (a, t, u) in zip(us, got):
        a1, t1, u1 = calc_attractor_single(u0, tr.flags, AFLAGS, tr.nstep, 1, 12)
        print(a, t, a1, t1, flush=True)
        assert (a, t) == (a1, t1)
        assert np.array_equal(u, u1)
        kinds.add(a)
    assert E.LAMINAR in kinds and E.TURBULENT in kinds  # both verdicts exercised


def test_bisection_round_matches_sequential(gpu):
    fl = cfa.DNSFlags(**FLAGS)
    tr = E.EdgeTracker(NX, NY, NZ, LX_C, LZ_C, fl, AFLAGS, tMin=1, tMax=12)
    # blends of norm about 0.01 j: the first ones decay, later ones do not
    base = random_field(NX, NY, NZ, LX_C, LZ_C, norm=0.05, decay=0.4, seed=800)
    uL, uH = 0.0 * base, base.copy()
    L, H, d = tr.bisection_round(uL, uH, nTraj=4)
    # sequential: blends j = 1, 2, ... until one is not laminar
    step = 1.0 / 5
    j, a = 0, E.LAMINAR
    while a == E.LAMINAR and j <= 4 + 10:
        j += 1
        al = j * step
        uT = uL.copy()
        uT *= (1.0 - al) / al
        uT += uH
        uT *= al
        a = calc_attractor_single(uT, tr.flags, AFLAGS, tr.nstep, 1, 12)[0]
    print("first non-laminar blend", j, a, flush=True)
    assert np.array_equal(H, uT)
    if j >= 3:
        al = (j - 2) * step
        expL = uL.copy()
        expL *= (1.0 - al) / al
        expL += uH
        expL *= al
        assert np.array_equal(L, expL)
    else:
        assert np.array_equal(L, uL)
    assert d == E.rel_l2_diff(L, H)


def calc_attractor_single_vardt(u0, flags, aflags, tMin, tMax):
    """CalcAttractor with its own DNS and TimeStep (edgetracking.cpp:767-812):
    after every time unit the CFL number decides a reset_dt."""
    from channelflow_amd.timestep import TimeStep
    dt = TimeStep(flags=flags)
    fu = cfa.FlowField(NX, NY, NZ, 3, LX_C, LZ_C, data=u0)
    fq = cfa.FlowField(NX, NY, NZ, 1, LX_C, LZ_C)
    cfa.PressureSolver(fu, flags=flags).solve(fq, fu)
    dns = cfa.DNS([fu, fq], flags.copy(dt=dt.dt()))
    u = np.array(u0, dtype=np.complex128)
    result, t, tDecay, dts = E.NO_ATTRACTOR, 0, 0, []
    while result == E.NO_ATTRACTOR and t < tMax:
        t += 1
        lastEcf = E.ecf(u)
        dns.advance_resident(dt.n())
        if dt.adjust(dns.CFL()):
            dns.reset_dt(dt.dt())
        dts.append(dt.dt())
        u = dns.state()[0]
        if t > tMin:
            result, tDecay = E.classify(u, tDecay, lastEcf, aflags)
    return result, t, u, dts


def test_attractors_variable_dt_match_single_runs(gpu):
    """flags.variabledt: each trajectory of the batched call adapts its own dt
    (its own step count per time unit) and ends bit for bit where its own DNS
    with its own TimeStep ends."""
    fl = cfa.DNSFlags(**dict(FLAGS, variabledt=True, dtmax=0.05))
    tr = E.EdgeTracker(NX, NY, NZ, LX_C, LZ_C, fl, AFLAGS, tMin=1, tMax=12)
    us = members()
    got = tr.attractors(us)
    changed = False
    for u0, (a, t, u) in zip(us, got):
        a1, t1, u1, dts = calc_attractor_single_vardt(u0, tr.flags, AFLAGS, 1, 12)
        print(a, t, a1, t1, "dt sequence", sorted(set(dts)), flush=True)
        changed = changed or len(set(dts)) > 1 or dts[0] != FLAGS["dt"]
        assert (a, t) == (a1, t1)
        assert np.array_equal(u, u1)
    assert changed  # dt did adapt
