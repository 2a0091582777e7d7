"""Batched and served integrations on the GPU.

Every member of a batch, and every request of the server, must give the same
result as its own single-field DNS (the per-member kernels never mix
members), and that result must stay within the 20-step bar of the oracle
(relative L2 <= 1e-10 per component). A member's result must not depend on
what else is in the batch or when it was admitted.
"""
import numpy as np
import pytest

import cfref as R
import channelflow_amd as cfa
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu

LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5

COUETTE = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
               nu=1 / 400, dt=0.02, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric",
               dealiasing="DealiasXZ")
POISEUILLE = dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 / 2000, nu=1 / 2000, dt=0.01,
                  timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational", dealiasing="DealiasXZ")


def case(i):
    """Member i of the mixed test batch: grid, flags, initial field."""
    grids = [(16, 17, 16), (24, 25, 24), (16, 17, 16), (24, 25, 24), (16, 33, 16), (16, 17, 16)]
    Nx, Ny, Nz = grids[i % len(grids)]
    fl = dict(COUETTE if i % 2 == 0 else POISEUILLE)
    if i == 5:
        fl.update(nu=1 / 300)  # same grid class, different tau data
    if i == 4:
        fl.update(timestepping="SBDF2")
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.05 + 0.02 * i, decay=0.4, seed=100 + i)
    return (Nx, Ny, Nz), fl, u0


def fields(grid, u0):
    Nx, Ny, Nz = grid
    return [cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u0), cfa.FlowField(Nx, Ny, Nz, 1, LX_C, LZ_C)]


def single(grid, fl, u0, n):
    f = fields(grid, u0)
    d = cfa.DNS(f, cfa.DNSFlags(**fl))
    d.advance_resident(n)
    u, q = d.state()
    return u, q, d.time()


def test_graph_replay_matches_eager(gpu):
    """Replayed hipGraphs give bit-identical states to eager launches, also
    across a join and a removal mid-run (new launch sequences, new captures),
    with one graph per grid class and with one graph per tick for the whole
    batch; the whole-batch mode launches at most one graph per tick."""
    def run(graphs):
        b = cfa.Batch()
        b.set_graphs(graphs)
        ids = []
        for i in range(4):
            grid, fl, u0 = case(i)
            ids.append(b.add(fields(grid, u0), cfa.DNSFlags(**fl)))
        b.advance(9)
        grid, fl, u0 = case(5)
        late = b.add(fields(grid, u0), cfa.DNSFlags(**fl))
        b.advance(7)
        b.remove(ids[1])
        b.advance(20)  # long enough for the whole batch's launch sequences to repeat (period 6)
        keep = [ids[0], ids[2], ids[3], late]
        return {k: b.state(k) for k in keep}, b.graph_stats(), b.tick_stats(), b.size()[1]

    eager, st_e, tk_e, _ = run(False)
    for mode in (cfa.Batch.GRAPHS_PER_CLASS, cfa.Batch.GRAPHS_WHOLE):
        graphed, st_g, tk_g, ncls = run(mode)
        print(f"graph mode {mode}: captures/replays {st_g}, ticks/launches/host us {tk_g} ({ncls} classes); "
              f"eager {st_e}", flush=True)
        for k in eager:
            assert np.array_equal(eager[k][0], graphed[k][0]) and np.array_equal(eager[k][1], graphed[k][1])
        # start-up substeps and every change of membership are new launch
        # sequences; each repeating one is captured once and then replayed
        assert st_g[0] > 0 and st_g[1] > 0
        assert tk_g[0] == tk_e[0]
        if mode == cfa.Batch.GRAPHS_WHOLE:
            assert ncls > 1 and tk_g[1] <= tk_g[0]
    assert st_e == (0, 0) and tk_e[1] == 0


def test_batch_cfl_matches_single(gpu):
    """DNS::CFL of batch members equals their own single DNS's, bit for bit."""
    b = cfa.Batch()
    ids, dnss = [], []
    for i in (0, 1, 4):
        grid, fl, u0 = case(i)
        ids.append(b.add(fields(grid, u0), cfa.DNSFlags(**fl)))
        d = cfa.DNS(fields(grid, u0), cfa.DNSFlags(**fl))
        d.advance_resident(7)
        dnss.append(d)
    b.advance(7)
    for mid, d in zip(ids, dnss):
        print(f"member {mid}: batch CFL {b.CFL(mid):.15e} single {d.CFL():.15e}", flush=True)
        assert b.CFL(mid) == d.CFL()


def test_batch_matches_single_and_oracle(gpu):
    n = 20
    b = cfa.Batch()
    ids = []
    for i in range(6):
        grid, fl, u0 = case(i)
        ids.append(b.add(fields(grid, u0), cfa.DNSFlags(**fl)))
    assert b.size() == (6, 3)
    b.advance(n)
    for i, mid in enumerate(ids):
        grid, fl, u0 = case(i)
        ub, qb = b.state(mid)
        us, qs, ts = single(grid, fl, u0, n)
        print(f"member {i} grid {grid} {fl['nonlinearity']} {fl['timestepping']}: batch vs single max|diff| "
              f"{np.max(np.abs(ub - us)):.3e}", flush=True)
        assert np.array_equal(ub, us) and np.array_equal(qb, qs)
        assert b.time(mid) == pytest.approx(ts, abs=1e-14)
        if i < 2:  # the oracle is slow; two members of different physics suffice
            g = R.Grid(*grid, LX_C, LZ_C)
            ref = R.OracleDNS(g, R.make_flags(**fl), u0, g.zeros(1))
            ref.advance(n)
            ur, _ = ref.state()
            for c in range(3):
                err = np.sqrt(R.l2dist2(g, ub[c:c + 1], ur[c:c + 1]) / R.l2norm2(g, ur[c:c + 1]))
                print(f"  component {c}: rel L2 vs oracle {err:.3e}", flush=True)
                assert err <= 1e-10


def test_member_alone_and_in_mix(gpu):
    grid, fl, u0 = case(1)
    alone = cfa.Batch()
    a = alone.add(fields(grid, u0), cfa.DNSFlags(**fl))
    alone.advance(12)
    ua, qa = alone.state(a)
    mix = cfa.Batch()
    for i in (0, 2, 3):
        g2, f2, v2 = case(i)
        mix.add(fields(g2, v2), cfa.DNSFlags(**f2))
    m = mix.add(fields(grid, u0), cfa.DNSFlags(**fl))
    mix.advance(5)
    g3, f3, v3 = case(5)
    mix.add(fields(g3, v3), cfa.DNSFlags(**f3))  # joins mid-run, in its init steps while the others are not
    mix.advance(7)
    um, qm = mix.state(m)
    print("alone vs mix max|diff|", np.max(np.abs(ua - um)), flush=True)
    assert np.array_equal(ua, um) and np.array_equal(qa, qm)


def test_server_results_match_single(gpu):
    horizons = [7, 3, 11, 5, 9, 4]
    srv = cfa.Server(budget_bytes=8e6)  # about two requests at a time: admission must queue
    rid = {}
    for i, h in enumerate(horizons):
        grid, fl, u0 = case(i)
        rid[srv.submit(fields(grid, u0), cfa.DNSFlags(**fl), h)] = i
    results = {}
    max_running = 0
    for _ in range(10000):
        done, running, queued = srv.step()
        max_running = max(max_running, running)
        for r in done:
            results[r] = srv.result(r)
        if len(results) == len(horizons):
            break
    assert len(results) == len(horizons)
    assert max_running >= 2
    for r, (u, q, t) in results.items():
        i = rid[r]
        grid, fl, u0 = case(i)
        us, qs, ts = single(grid, fl, u0, horizons[i])
        print(f"request {i} horizon {horizons[i]}: served vs single max|diff| {np.max(np.abs(u - us)):.3e}",
              flush=True)
        assert np.array_equal(u, us) and np.array_equal(q, qs)
        assert t == pytest.approx(ts, abs=1e-14)


def test_table_arena_recycle_mid_plan(gpu, monkeypatch):
    """A tick whose launch tables fill the table arena part-way through its
    plan: the plan is redone in the fresh arena, so no record keeps an address
    that a later table of the same tick overwrote. With a 4 KiB arena every few
    ticks recycle it, in the middle of a plan of mixed forms (rotational and
    skew-symmetric members in one class); every member must equal its run
    under the default 4 MiB arena bit for bit, with graphs on."""
    def run():
        b = cfa.Batch()
        ids = []
        for i in range(4):
            fl = dict(COUETTE, nonlinearity="Rotational" if i % 2 else "SkewSymmetric")
            u0 = random_field(16, 17, 16, LX_C, LZ_C, norm=0.05 + 0.02 * i, decay=0.4, seed=300 + i)
            ids.append(b.add(fields((16, 17, 16), u0), cfa.DNSFlags(**fl)))
        b.advance(12)
        return [b.state(k) for k in ids], b.graph_stats()

    ref, st_ref = run()
    monkeypatch.setenv("CFGPU_TABLE_ARENA_BYTES", "4096")
    small, st_small = run()
    print("graph captures/replays: 4 MiB arena", st_ref, "4 KiB arena", st_small, flush=True)
    for (ur, qr), (us, qs) in zip(ref, small):
        assert np.array_equal(ur, us) and np.array_equal(qr, qs)


def test_zpass_staged_and_unstaged_agree(gpu, monkeypatch):
    """The wavefront z-pass has two forms: one that stages each transform's
    input fields in LDS ahead of use (LDS DMA copies), the default at Nz = 64,
    and one that loads them as it goes. Their arithmetic is the same up to
    which product of a*b + c*d the compiler fuses into a multiply-add, so every
    member must agree under the two to rounding: after 8 steps 1e-14 of the
    velocity's size, and 1e-12 of the pressure's, which the tau correction
    makes sensitive to rounding (a misplaced or stale staged bin would be off
    by the field's own size, 1e-3 and more). Members cover the four
    nonlinear forms, the wavefront lengths 16, 32, 48 and 64, the rotational
    form at 128 (whose staged kernel is compiled for 2 waves per SIMD), and
    grids without dealiasing (Nz/2 live kz bins, the staging's largest)."""
    specs = [((16, 17, 16), "SkewSymmetric", "DealiasXZ"), ((32, 17, 32), "Rotational", "DealiasXZ"),
             ((48, 17, 48), "Convection", "DealiasXZ"), ((64, 17, 64), "SkewSymmetric", "DealiasXZ"),
             ((32, 17, 32), "Divergence", "NoDealiasing"), ((48, 17, 48), "SkewSymmetric", "NoDealiasing"),
             ((64, 17, 64), "Rotational", "DealiasXZ"), ((64, 17, 64), "Convection", "NoDealiasing"),
             ((128, 17, 128), "Rotational", "DealiasXZ")]

    def run():
        b = cfa.Batch()
        ids = []
        for i, (grid, nl, de) in enumerate(specs):
            fl = dict(COUETTE, nonlinearity=nl, dealiasing=de)
            u0 = random_field(*grid, LX_C, LZ_C, norm=0.05 + 0.02 * i, decay=0.4, seed=400 + i)
            ids.append(b.add(fields(grid, u0), cfa.DNSFlags(**fl)))
        b.advance(8)
        return [b.state(k) for k in ids]

    monkeypatch.setenv("CFGPU_ZWAVE_STAGED", "0")
    loads = run()
    monkeypatch.setenv("CFGPU_ZWAVE_STAGED", "1")
    staged = run()
    for (grid, nl, de), (ul, ql), (us, qs) in zip(specs, loads, staged):
        du = float(np.max(np.abs(ul - us))) / float(np.max(np.abs(ul)))
        dq = float(np.max(np.abs(ql - qs))) / float(np.max(np.abs(ql)))
        print(grid, nl, de, f"max|staged - unstaged| / max|u| {du:.2e}, q {dq:.2e}", flush=True)
        assert du <= 1e-12 and dq <= 1e-9


def _device_free_bytes():
    """hipMemGetInfo through the HIP runtime the library already loaded."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_server_budget_covers_device_memory(gpu):
    """The memory the server counts against its budget (pool blocks at
    hipMalloc's rounding, work buffers, tau sets) grows by at least the device
    memory its integrations really take: state, histories, start-up buffers,
    work buffers and tau data (a distinct nu per request, so no tau set is
    shared). Measured with hipMemGetInfo around the admission and first tick of
    six requests on one grid class that already exists."""
    grid = (32, 33, 32)
    Nx, Ny, Nz = grid
    srv = cfa.Server(budget_bytes=64e9)

    def submit(k):
        fl = dict(COUETTE, nu=1.0 / (400 + 7 * k))
        u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.05, decay=0.4, seed=300 + k)
        srv.submit(fields(grid, u0), cfa.DNSFlags(**fl), 50)

    submit(0)
    srv.step()  # the class exists: geometry, tables, first work buffers
    used0, free0 = srv.memory(), _device_free_bytes()
    for k in range(1, 7):
        submit(k)
    srv.step()  # admits the six and runs their first tick (start-up buffers, tau data)
    used1, free1 = srv.memory(), _device_free_bytes()
    estimate, actual = used1 - used0, free0 - free1
    print(f"six 32^3 requests: memory counted {estimate / 1e6:.1f} MB, device memory taken {actual / 1e6:.1f} MB",
          flush=True)
    assert actual > 0 and actual <= estimate + (2 << 20)  # hipMemGetInfo moves in 2 MiB steps


def test_server_backfills_past_a_large_request(gpu):
    """A request that does not fit the free budget yet is passed by smaller
    ones that do (backfilling), and is admitted once memory frees up."""
    small, big = (16, 17, 16), (24, 25, 24)
    # admission needs (pool blocks at hipMalloc's rounding, 4 tau sets, work
    # buffer growth): a small request 1.7 MB (2.5 MB for the third of its
    # class), the large one 6.3 MB
    srv = cfa.Server(budget_bytes=7.0e6)
    ids = []
    for i, g in enumerate([small, big, small, small]):
        fl = dict(COUETTE)
        u0 = random_field(*g, LX_C, LZ_C, norm=0.05, decay=0.4, seed=400 + i)
        ids.append(srv.submit(fields(g, u0), cfa.DNSFlags(**fl), 6 if g == small else 4))
    order, running_max, held_max = [], 0, 0.0
    for _ in range(1000):
        done, running, queued = srv.step()
        running_max = max(running_max, running)
        held_max = max(held_max, srv.memory())
        for r in done:
            srv.result(r)
            order.append(r)
        if len(order) == 4:
            break
    print(f"completion order {order} (submitted {ids}), most running at once {running_max}, "
          f"most memory held {held_max / 1e6:.2f} MB", flush=True)
    assert len(order) == 4
    assert order.index(ids[2]) < order.index(ids[1])  # a small request passed the large one
    assert held_max <= 7.0e6
