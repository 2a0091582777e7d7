"""The benchmark configurations at their real grid sizes on the HIP path,
against the oracle (relative L2 <= 1e-10 per component after 20 steps) and,
for batching and serving, against the plain single-field path bit for bit.

  * Batched ensemble: two members of each grid class (32x33x32, 48x49x48,
    64x65x64) with the ensemble's flags and log-uniform norms, in one Batch.
  * Multiple shooting: 2 shots and 1 Jacobian direction at 32x49x32, G and the
    finite-difference product against the oracle's.
  * Serving: one request per grid class 32^3 ... 128^3, plane Couette
    (skew-symmetric, bulk velocity) and plane Poiseuille (rotational,
    pressure gradient) alternating, through the Server with a join mid-run:
    served == single bit for bit; 32^3 ... 64^3 against the oracle, 96^3 and
    128^3 against committed oracle fixtures (tests/golden/make_c5_reference.py).
"""
import json
import os

import numpy as np
import pytest

import bench
import cfref as R
import channelflow_amd as cfa
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LX, LZ = bench.LX_C, bench.LZ_C


def rel(g, a, b):
    return np.array([np.sqrt(R.l2dist2(g, a[i:i + 1], b[i:i + 1]) / R.l2norm2(g, b[i:i + 1]))
                     for i in range(a.shape[0])])


def flow_fields(grid, u):
    Nx, Ny, Nz = grid
    return [cfa.FlowField(Nx, Ny, Nz, 3, LX, LZ, data=u), cfa.FlowField(Nx, Ny, Nz, 1, LX, LZ)]


def oracle_run(grid, flags, u, n):
    g = R.Grid(*grid, LX, LZ)
    d = R.OracleDNS(g, R.make_flags(**flags), u, g.zeros(1))
    d.advance(n)
    return g, d.state()


def single_run(grid, flags, u, n):
    d = cfa.DNS(flow_fields(grid, u), cfa.DNSFlags(**flags))
    d.advance_resident(n)
    return d.state()


def test_ensemble_classes_batch_vs_oracle(gpu):
    members = bench.c3_members()
    pick = [0, 1, 24, 25, 48, 49]  # two members of each grid class
    b = cfa.Batch()
    ids = []
    for m in pick:
        mem = members[m]
        ids.append(b.add(flow_fields(mem["grid"], bench.member_field(mem)), cfa.DNSFlags(**mem["flags"])))
    assert b.size() == (6, 3)
    b.advance(20)
    for k, m in enumerate(pick):
        mem = members[m]
        u0 = bench.member_field(mem)
        print(f"member {m}: {mem['grid']} norm {mem['norm']:.3e}", flush=True)
        ub, qb = b.state(ids[k])
        g, (uo, qo) = oracle_run(mem["grid"], mem["flags"], u0, 20)
        e = rel(g, ub, uo)
        print("  20 steps, rel L2 per component vs oracle", e, flush=True)
        assert np.all(e <= 1e-10), e
        if k % 2 == 0:  # batch == plain single-field path, bit for bit
            us, qs = single_run(mem["grid"], mem["flags"], u0, 20)
            assert np.array_equal(ub, us) and np.array_equal(qb, qs)


def test_multishooting_c4_grid_vs_oracle(gpu):
    from channelflow_amd.dsi import MultiShooting
    Nx, Ny, Nz = 32, 49, 32
    T = 0.4  # 20 steps of dt = 0.02
    shots = [random_field(Nx, Ny, Nz, LX, LZ, norm=0.1 + 0.1 * i, decay=0.3, seed=2024, member=i) for i in range(2)]
    dirs = [[random_field(Nx, Ny, Nz, LX, LZ, norm=1.0, decay=0.3, seed=2024 + 7919, member=1000 * i)
             for i in range(2)]]
    flags = dict(bench.COUETTE, variabledt=False)
    ms = MultiShooting(Nx, Ny, Nz, LX, LZ, cfa.DNSFlags(**flags), T)
    G, DG, eps = ms.eval_with_jacobians(shots, dirs, epsDx=1e-7)
    assert ms.nsteps == 20
    of = dict(bench.COUETTE, dt=ms.flags.dt)
    n = len(shots)
    fo = [oracle_run((Nx, Ny, Nz), of, u, ms.nsteps)[1][0] for u in shots]
    g = R.Grid(Nx, Ny, Nz, LX, LZ)
    Go = [fo[i] - shots[(i + 1) % n] for i in range(n)]
    e = eps[0]
    xp = [shots[i] + e * dirs[0][i] for i in range(n)]
    fp = [oracle_run((Nx, Ny, Nz), of, u, ms.nsteps)[1][0] for u in xp]
    DGo = [(1 / e) * ((fp[i] - xp[(i + 1) % n]) - Go[i]) for i in range(n)]
    for i in range(n):
        ef = rel(g, G[i] + shots[(i + 1) % n], fo[i])
        ej = np.max(rel(g, DG[0][i], DGo[i]))
        print(f"shot {i}: f^T rel L2 per component {ef}, Jacobian product {ej:.2e}", flush=True)
        assert np.all(ef <= 1e-10) and ej <= 1e-6


def serving_cases():
    """One request per grid class; 96^3 and 128^3 are the fixtures' cases."""
    from golden.make_c5_reference import CASES
    cases = [dict(grid=(32, 33, 32), flags=dict(bench.COUETTE, nu=1 / 300), norm=0.3, seed=2024, member=32),
             dict(grid=(48, 49, 48), flags=bench.poiseuille(1 / 1500, 0.01 * 32 / 48), norm=0.1, seed=2024,
                  member=48),
             dict(grid=(64, 65, 64), flags=dict(bench.COUETTE, nu=1 / 500, dt=0.01), norm=0.01, seed=2024,
                  member=64)]
    for name in ("c5_96_poiseuille", "c5_128_couette"):
        c = CASES[name]
        cases.append(dict(grid=c["grid"], flags=c["flags"], norm=c["norm"], seed=c["seed"], member=c["member"],
                          fixture=name))
    for c in cases:
        Nx, Ny, Nz = c["grid"]
        c["u0"] = random_field(Nx, Ny, Nz, LX, LZ, norm=c["norm"], decay=0.3, seed=c["seed"], member=c["member"])
    return cases


def test_serving_classes_vs_single_and_oracle(gpu):
    steps = 20
    cases = serving_cases()
    srv = cfa.Server(64e9)
    rid = {}
    for k, c in enumerate(cases[:3]):
        rid[srv.submit(flow_fields(c["grid"], c["u0"]), cfa.DNSFlags(**c["flags"]), steps)] = k
    results = {}
    ticks = 0
    while len(results) < len(cases):
        if ticks == 7:  # the two large requests join while the small ones run
            for k in (3, 4):
                c = cases[k]
                rid[srv.submit(flow_fields(c["grid"], c["u0"]), cfa.DNSFlags(**c["flags"]), steps)] = k
        done, running, queued = srv.step()
        ticks += 1
        for r in done:
            results[rid[r]] = srv.result(r)
        assert ticks < 10000
    for k, c in enumerate(cases):
        print(f"request {k}: {c['grid']} {c['flags']['nonlinearity']}", flush=True)
        us, qs, ts = results[k]
        ug, qg = single_run(c["grid"], c["flags"], c["u0"], steps)
        assert np.array_equal(us, ug) and np.array_equal(qs, qg)
        Nx, Ny, Nz = c["grid"]
        g = R.Grid(Nx, Ny, Nz, LX, LZ)
        if "fixture" not in c:
            _, (uo, _) = oracle_run(c["grid"], c["flags"], c["u0"], steps)
            e = rel(g, us, uo)
            print("  served == single; rel L2 per component vs oracle", e, flush=True)
            assert np.all(e <= 1e-10), e
        else:
            meta = json.load(open(os.path.join(HERE, "golden", c["fixture"] + ".json")))
            ref = np.fromfile(os.path.join(HERE, "golden", c["fixture"] + "_modes.c128"),
                              dtype=np.complex128).reshape(meta["shape"])
            norms = np.array([np.sqrt(R.l2norm2(g, us[i:i + 1])) for i in range(3)])
            en = np.abs(norms - np.array(meta["l2norm_u"])) / np.array(meta["l2norm_u"])
            got = np.stack([us[:, :, kx % Nx, kz] for kx, kz in meta["modes"]], axis=0)
            em = np.array([np.linalg.norm(got[:, i] - ref[:, i]) / np.linalg.norm(ref[:, i]) for i in range(3)])
            print("  served == single; vs oracle fixture: norm rel diff", en, "sampled modes", em, flush=True)
            assert np.all(en <= 1e-10) and np.all(em <= 1e-10)
            assert abs(ts - meta["time"]) < 1e-12
