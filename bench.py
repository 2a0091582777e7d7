"""Benchmark of the MI355X DNS time step (one GPU).

Default workload (the headline): a batched ensemble of 64 independent plane
Couette integrations at Re = 400 on uneven grids -- 24 at 32x33x32, 24 at
48x49x48, 16 at 64x65x64, box Lx = 2 pi/1.14, Lz = 2 pi/2.5, FP64,
skew-symmetric nonlinearity, bulk velocity, 2/3 dealiasing, SBDF3 with
dt = 0.02 and the SMRK2 start. Member m starts from the seeded
divergence-free perturbation (coefficients scaled by 0.3^(|kx|+|kz|+ny/4))
with an L2 norm drawn log-uniform in [1e-3, 0.3] from seed 2024 + m. A step
is one DNS::advance step of every member; states, histories and tau data are
resident in HBM before the timed region starts.

The same JSON line also carries
  plain_gpu   the plain single-field GPU path on the same members (one
              integration at a time) and the batched/plain ratio;
  served      a serving run: Poisson arrivals of Couette (Re 300-500) and
              Poiseuille (Re_bulk 1500-3000) requests on grids 32^3 ... 128^3,
              horizons 100-2000 steps, admitted by HBM budget, with p50/p99
              completion latency and aggregate grid-point*steps/s, beside the
              same requests run one at a time on the plain path; `saturated`
              repeats it with every request queued at t = 0 (the capacity),
              `budget_4GB` the same under a 4 GB budget that binds
              (admission with backfilling decides what runs at once);
  multishooting one call of the multiple-shooting batch client: the residual
              of 16 shots and 8 Jacobian products, 144 integrations of
              32x49x32 over T = 4 in one Batch, timed whole, beside the same
              integrations on the plain path;
  single_c2   the plain 192x129x192 Poiseuille channel (rotational, SBDF3,
              dt = 0.005) and its per-kernel times;
  graphs      hipGraph captures and replays inside the timed region (one
              graph per grid class and tick, the default), graph launches
              and host microseconds per tick, and the same steps with one
              graph per tick for the whole batch (`whole_batch`);
  roofline    the dominant kernel of the headline workload (HIP events over
              a profiled pass of the same members right after the timed
              region, one grid class at a time so that a launch's events
              time that kernel alone; replay is off while events are
              recorded);
  cpu_baseline the oracle (oracle/cpu_ref.cpp) on a bounded sample: one
              member per grid class on one core, and one member per core on
              all host cores, with the CPU model.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
LX_C, LZ_C = 2 * np.pi / 1.14, 2 * np.pi / 2.5
METRIC = "grid-pt·steps/s batched+served vs plain GPU at 1e-10 rel-L2; p50/p99 latency"

COUETTE = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
               nu=1.0 / 400, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric",
               dealiasing="DealiasXZ", taucorrection=True, dt=0.02)


def poiseuille(nu, dt):
    """Laminar Poiseuille with bulk velocity 1 (centreline 1.5): dPdx = -3 nu."""
    return dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 * nu, nu=nu, timestepping="SBDF3",
                initstepping="SMRK2", nonlinearity="Rotational", dealiasing="DealiasXZ", taucorrection=True, dt=dt)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def c3_members(seed=2024):
    """64 Couette members on three grids, norms log-uniform in [1e-3, 0.3]."""
    grids = [(32, 33, 32)] * 24 + [(48, 49, 48)] * 24 + [(64, 65, 64)] * 16
    out = []
    for m, g in enumerate(grids):
        r = np.random.default_rng(seed + m)
        norm = math.exp(r.uniform(math.log(1e-3), math.log(0.3)))
        out.append(dict(grid=g, flags=COUETTE, norm=norm, decay=0.3, seed=seed, member=m))
    return out


def c4_members(seed=2024):
    """16 base fields at 32x49x32 (norm U[0.1, 0.2]), each with 8 copies
    perturbed by 1e-7 times a unit-norm divergence-free field: 144 members."""
    out = []
    for b in range(16):
        r = np.random.default_rng(seed + b)
        norm = r.uniform(0.1, 0.2)
        out.append(dict(grid=(32, 49, 32), flags=COUETTE, norm=norm, decay=0.3, seed=seed, member=b))
        for k in range(8):
            out.append(dict(grid=(32, 49, 32), flags=COUETTE, norm=norm, decay=0.3, seed=seed, member=b,
                            delta=(seed + 7919, 1000 * b + k)))
    return out


def run_multishooting(cfa, T=4.0, nshot=16, ndir=8, seed=2024, plain_sample=9):
    """One call of the multiple-shooting batch client (channelflow_amd.dsi):
    G(x) of 16 shots plus 8 forward-difference Jacobian products at
    epsDx = 1e-7, i.e. 16 x (1 + 8) = 144 integrations of 32x49x32 over T,
    in one Batch. Timed whole (member set-up, tau data, autotuning, the steps,
    results copied back), as a Newton-Krylov iteration would pay it. The plain
    path runs `plain_sample` of the same integrations one DNS at a time (one
    shot and its 8 perturbed copies) and is scaled to 144."""
    from channelflow_amd.dsi import MultiShooting
    from channelflow_amd.fields import random_field
    Nx, Ny, Nz = 32, 49, 32
    shots = [member_field(m) for m in c4_members(seed) if "delta" not in m][:nshot]
    dirs = [[random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=1.0, decay=0.3, seed=seed + 7919, member=1000 * b + k)
             for b in range(nshot)] for k in range(ndir)]
    fl = cfa.DNSFlags(**dict(COUETTE, variabledt=False))
    ms = MultiShooting(Nx, Ny, Nz, LX_C, LZ_C, fl, T)
    t0 = time.perf_counter()
    G, DG, eps = ms.eval_with_jacobians(shots, dirs, epsDx=1e-7)
    wall = time.perf_counter() - t0
    n_int = nshot * (1 + ndir)
    pts = Nx * Ny * Nz
    sample = [shots[0]] + [shots[0] + eps[k] * dirs[k][0] for k in range(ndir)]
    sample = sample[:plain_sample]
    t0 = time.perf_counter()
    for u in sample:
        d = cfa.DNS(flow_fields(cfa, (Nx, Ny, Nz), u), ms.flags)
        d.advance_resident(ms.nsteps)
        d.state()
    wall_p = (time.perf_counter() - t0) * n_int / len(sample)
    gnorm = math.sqrt(sum(float(np.vdot(g, g).real) for g in G))
    return {"workload": f"C4 multi-shooting call: {nshot} shots x (1 + {ndir} Jacobian directions) = {n_int} "
                        f"plane Couette Re=400 integrations at {Nx}x{Ny}x{Nz}, T={T} ({ms.nsteps} SBDF3 steps, "
                        f"dt={ms.flags.dt}), epsDx=1e-7",
            "integrations": n_int, "seconds": wall, "value": n_int * pts * ms.nsteps / wall,
            "unit": "grid-pt*steps/s", "plain_gpu_seconds": wall_p,
            "plain_sample": f"{len(sample)} of the {n_int} integrations one DNS at a time, scaled",
            "batched_over_plain": wall_p / wall, "residual_coefficient_norm": gnorm}


def member_field(mem, Lx=LX_C, Lz=LZ_C):
    from channelflow_amd.fields import random_field
    Nx, Ny, Nz = mem["grid"]
    u = random_field(Nx, Ny, Nz, Lx, Lz, norm=mem["norm"], decay=mem["decay"], seed=mem["seed"],
                     member=mem["member"])
    if "delta" in mem:
        s, k = mem["delta"]
        u = u + 1e-7 * random_field(Nx, Ny, Nz, Lx, Lz, norm=1.0, decay=mem["decay"], seed=s, member=k)
    return u


def flow_fields(cfa, grid, u):
    Nx, Ny, Nz = grid
    return [cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u), cfa.FlowField(Nx, Ny, Nz, 1, LX_C, LZ_C)]


def live(grid):
    Nx, Ny, Nz = grid
    return Ny * (2 * (Nx // 3) - 1) * (Nz // 3)  # complex values per component


WAVE_FFT_LENGTHS = (16, 32, 48, 64, 96, 128, 192, 256)  # x/z lengths of the wavefront-local FFT kernels


def kernel_bytes(grid, method):
    """Compulsory HBM bytes of one member per launch of each kernel (inputs
    read once, outputs written once), SBDF3 main steps."""
    Nx, Ny, Nz = grid
    kxd, kzd = Nx // 3 - 1, Nz // 3 - 1
    M = (2 * kxd + 1) * (kzd + 1)
    L = 16.0 * Ny * M                     # one complex component array in the compact spectral layout
    X = 16.0 * Ny * Nx * (kzd + 1)        # one complex field, x physical, kz spectral
    nyout = {"Rotational": 6, "SkewSymmetric": 6, "Convection": 6, "Divergence": 3}[method]
    nxout = {"Rotational": 6, "SkewSymmetric": 9, "Convection": 9, "Divergence": 3}[method]
    nzout = {"Rotational": 3, "SkewSymmetric": 8, "Convection": 3, "Divergence": 6}[method]
    # the skew-symmetric forward x-FFT of the wavefront lengths adds the
    # x-derivative part of the divergence and writes f_c and u_c u_1: 6 fields
    nfo = 6 if method == "SkewSymmetric" and Nx in WAVE_FFT_LENGTHS else nzout
    return {
        "k_yinv": (3 + nyout) * L,
        "k_xinv": nyout * L + nxout * X,
        "k_zpass": nxout * X + nzout * X,
        "k_xfwd": nzout * X + nfo * L,
        "k_yfwd": nfo * L + 3 * L,
        # SBDF3 right-hand side u_n..u_{n-2}, f_n..f_{n-2} read, u (3) and q (1)
        # written: 22 L. The per-mode tau factors are shared by every member of a
        # class and are overhead (tau_factor_bytes), not algorithmic bytes.
        "k_tau_solve": (18 + 4) * L,
    }


def tau_factor_bytes(grid):
    """Bytes of one tau set (one lambda) of a grid class: 12 Ny-vectors of
    doubles and 6 scalars per live mode, read by every tau launch of the
    class whatever its member count."""
    Nx, Ny, Nz = grid
    M = (2 * (Nx // 3 - 1) + 1) * (Nz // 3)
    return 8.0 * (12 * Ny * M + 6 * M)


# ---------------------------------------------------------------------------
# runs
# ---------------------------------------------------------------------------
class FiniteAdvance:
    """batch.advance(n) for an ensemble whose members may stop being finite.
    Such members are stopped, as the reference's batch client stops them
    (channelflow/cfdsi.cpp:769-774): removed from the batch and recorded in
    `dropped` with the step count after which they were found; `ids` (member
    index -> batch id) keeps the others."""

    def __init__(self, cfa, batch, members):
        self.cfa, self.batch, self.members = cfa, batch, members
        self.ids, self.dropped, self.steps = {}, [], 0

    def add(self, i):
        m = self.members[i]
        self.ids[i] = self.batch.add(flow_fields(self.cfa, m["grid"], member_field(m)),
                                     self.cfa.DNSFlags(**m["flags"]))

    def points(self):
        return sum(np.prod(self.members[i]["grid"]) for i in self.ids)

    def __call__(self, nsteps):
        try:
            self.batch.advance(nsteps)
            bad = []
        except self.cfa.CfgError as e:
            if e.code != self.cfa.ERR_NONFINITE:
                raise
            bad = [i for i, k in self.ids.items() if not self.batch.finite(k)]
        self.steps += nsteps
        for i in bad:
            self.batch.remove(self.ids.pop(i))
            m = self.members[i]
            self.dropped.append(dict(member=i, grid="x".join(map(str, m["grid"])), norm=round(m["norm"], 4),
                                     nonfinite_after_step=self.steps))
        return bad


def run_batched(cfa, members, steps, warmup, prof_steps):
    """Timed region: `steps` steps of the whole ensemble with each class's
    repeating ticks replayed as hipGraphs and no events between launches.
    Then the per-kernel durations: each grid class alone (its members in a
    Batch of their own, warmed up), `prof_steps` steps with HIP events around
    every launch (graphs off while events are on). One class at a time, so a
    launch's events time that kernel, not the other classes' kernels running
    beside it on their streams; the sums over the classes are returned."""
    batch = cfa.Batch()
    adv = FiniteAdvance(cfa, batch, members)
    for i in range(len(members)):
        adv.add(i)
    adv(warmup)                           # SMRK2 start + warm-up SBDF3 steps (and graph captures), untimed
    extra = steady_state(batch, batch.size()[1], advance=adv)
    c0, r0 = batch.graph_stats()
    k0, l0, h0 = batch.tick_stats()
    t0 = time.perf_counter()
    # returns after every stream has drained; a member found not finite at
    # the end has run every step, but its points are not counted
    adv(steps)
    wall = time.perf_counter() - t0
    pts = adv.points()
    timed = sorted(adv.ids)  # the members counted in the timed region
    c1, r1 = batch.graph_stats()
    k1, l1, h1 = batch.tick_stats()
    # the same steps with one graph per tick for the whole batch, the grid
    # classes as its parallel branches
    batch.set_graphs(cfa.Batch.GRAPHS_WHOLE)
    steady_state(batch, 1, advance=adv)
    k2, l2, h2 = batch.tick_stats()
    t2 = time.perf_counter()
    adv(steps)
    wall_w = time.perf_counter() - t2
    pts_w = adv.points()
    k3, l3, h3 = batch.tick_stats()
    batch.set_graphs(True)
    kt = {}
    for g in sorted({m["grid"] for m in members}):
        cls = cfa.Batch()
        for mem in members:
            if mem["grid"] == g:
                cls.add(flow_fields(cfa, mem["grid"], member_field(mem)), cfa.DNSFlags(**mem["flags"]))
        cls.advance(warmup)
        cls.profile(True)
        cls.advance(prof_steps)
        for k, (ms, n) in cls.kernel_times().items():
            a = kt.setdefault(k, [0.0, 0])
            a[0] += ms
            a[1] += n
        cls.profile(False)
        del cls
    kt = {k: tuple(v) for k, v in kt.items()}
    for i, k in adv.ids.items():  # every member still in the batch is finite (advance checked it)
        if not batch.finite(k):
            raise SystemExit(f"non-finite state of member {i}")
    return dict(value=pts * steps / wall, wall=wall, kt=kt, pts=pts, classes=batch.size()[1],
                dropped_nonfinite=adv.dropped, steps_run=adv.steps, timed_members=timed,
                graphs={"mode": "one graph per grid class and tick, the classes on their own streams (the "
                                "default)",
                        "captured_in_timed_region": c1 - c0, "replayed_in_timed_region": r1 - r0,
                        "ticks_in_timed_region": k1 - k0, "graph_launches_per_step": (l1 - l0) / steps,
                        "host_us_per_tick": (h1 - h0) / max(1, k1 - k0), "warmup_steps_added": extra,
                        "whole_batch": {"mode": "one graph per tick for the whole batch, the grid classes as "
                                                "parallel branches; same members and steps, after the timed region",
                                        "value": pts_w * steps / wall_w, "ms_per_step": wall_w / steps * 1e3,
                                        "graph_launches_per_step": (l3 - l2) / steps,
                                        "host_us_per_tick": (h3 - h2) / max(1, k3 - k2)}})


def steady_state(batch, per_tick, period=6, max_extra=60, advance=None):
    """Extends the warm-up until the ensemble's ticks are all graph replays:
    a block of `period` steps (a multiple of the SBDF3 history rotation) that
    captures nothing and replays `per_tick` graphs per tick (the grid classes
    with one graph per class, 1 with one graph for the whole batch). Steps
    through `advance` (default batch.advance). Returns the steps added."""
    advance = advance or batch.advance
    extra = 0
    while extra < max_extra:
        c0, r0 = batch.graph_stats()
        k0 = batch.tick_stats()[0]
        advance(period)
        extra += period
        c1, r1 = batch.graph_stats()
        k1 = batch.tick_stats()[0]
        if c1 == c0 and k1 - k0 == period and r1 - r0 == per_tick * period:
            return extra
    raise SystemExit(f"bench: no graph steady state after {max_extra} extra warm-up steps")


def run_plain(cfa, members, steps, warmup):
    """The plain single-field GPU path: one DNS per integration, one at a
    time. Each grid class is timed on its first member; the ensemble's plain
    time is the sum over members."""
    per_class = {}
    for mem in members:
        key = (mem["grid"], tuple(sorted(mem["flags"].items())))
        if key in per_class:
            continue
        dns = cfa.DNS(flow_fields(cfa, mem["grid"], member_field(mem)), cfa.DNSFlags(**mem["flags"]))
        dns.advance_resident(warmup)
        t0 = time.perf_counter()
        dns.advance_resident(steps)
        per_class[key] = (time.perf_counter() - t0) / steps
    t_plain = sum(per_class[(m["grid"], tuple(sorted(m["flags"].items())))] for m in members)
    pts = sum(np.prod(m["grid"]) for m in members)
    return dict(value=pts / t_plain, ms_per_ensemble_step=t_plain * 1e3,
                ms_per_step_by_grid={"x".join(map(str, k[0])): round(v * 1e3, 4) for k, v in per_class.items()})


def serve_requests(n, seed):
    """C5 request stream: request descriptions. The time step shrinks with the
    grid (dt = dt_32 * 32 / Nx) so that the explicit advection number
    kx_max U_max dt stays at its 32^3 value (0.23 Couette, 0.17 Poiseuille):
    a fixed dt = 0.02 puts 128^3 Couette past the SBDF3 stability limit."""
    r = np.random.default_rng(seed)
    grids = [(32, 33, 32), (48, 49, 48), (64, 65, 64), (96, 97, 96), (128, 129, 128)]
    reqs = []
    for i in range(n):
        g = grids[r.integers(len(grids))]
        if r.uniform() < 0.5:
            fl = dict(COUETTE, nu=1.0 / r.uniform(300, 500), dt=0.02 * 32 / g[0])
        else:
            fl = poiseuille(1.0 / r.uniform(1500, 3000), 0.01 * 32 / g[0])
        reqs.append(dict(grid=g, flags=fl, norm=math.exp(r.uniform(math.log(1e-3), math.log(0.3))), decay=0.3,
                         seed=seed, member=10000 + i, horizon=int(r.integers(100, 2001))))
    return reqs


def run_served(cfa, n, rate, seed, budget):
    """The request stream through a Server: Poisson arrivals at `rate` per
    second, or, with rate None, every request queued at t = 0 (the server's
    capacity, with latency then counting the time spent queued)."""
    reqs = serve_requests(n, seed)
    if rate is None:
        arrivals = np.zeros(n)
    else:
        arrivals = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, n))
    fields = [flow_fields(cfa, q["grid"], member_field(q)) for q in reqs]
    srv = cfa.Server(budget)
    rid_of = {}
    submitted = 0
    lat = {}
    t0 = time.perf_counter()
    max_running = 0
    max_used = 0.0
    next_report = 5.0
    while len(lat) < n:
        now = time.perf_counter() - t0
        while submitted < n and arrivals[submitted] <= now:
            q = reqs[submitted]
            rid_of[srv.submit(fields[submitted], cfa.DNSFlags(**q["flags"]), q["horizon"])] = submitted
            submitted += 1
        done, running, queued = srv.step()
        max_running = max(max_running, running)
        max_used = max(max_used, srv.memory())
        now = time.perf_counter() - t0
        for rid in done:
            i = rid_of[rid]
            u, _, _ = srv.result(rid)
            if not np.all(np.isfinite(u)):
                raise SystemExit(f"non-finite served result {i}")
            lat[i] = now - arrivals[i]
        if now > next_report:
            print(f"bench: serving t={now:.0f}s submitted {submitted}/{n} done {len(lat)} running {running} "
                  f"queued {queued}", file=sys.stderr, flush=True)
            next_report += 5.0
        if running == 0 and queued == 0 and submitted < n:
            time.sleep(max(0.0, min(0.01, arrivals[submitted] - now)))
    wall = time.perf_counter() - t0
    work = sum(np.prod(q["grid"]) * q["horizon"] for q in reqs)
    lats = np.array([lat[i] for i in range(n)])
    return dict(value=work / wall, unit="grid-pt*steps/s", requests=n, arrival_rate_per_s=rate,
                p50_latency_s=float(np.percentile(lats, 50)), p99_latency_s=float(np.percentile(lats, 99)),
                wall_s=wall, max_concurrent=max_running, hbm_budget_bytes=budget, max_admitted_bytes=max_used,
                work_grid_pt_steps=float(work))


def plain_for_requests(cfa, n, seed, steps, warmup):
    """The served request mix on the plain single-field path: requests one
    after another, each at the per-step time of its (grid, flow) class alone."""
    reqs = serve_requests(n, seed)
    per = {}
    for q in reqs:
        key = (q["grid"], q["flags"]["nonlinearity"])
        if key in per:
            continue
        dns = cfa.DNS(flow_fields(cfa, q["grid"], member_field(q)), cfa.DNSFlags(**q["flags"]))
        dns.advance_resident(warmup)
        t0 = time.perf_counter()
        dns.advance_resident(steps)
        per[key] = (time.perf_counter() - t0) / steps
    t = sum(q["horizon"] * per[(q["grid"], q["flags"]["nonlinearity"])] for q in reqs)
    work = sum(np.prod(q["grid"]) * q["horizon"] for q in reqs)
    return dict(value=work / t, seconds=t,
                ms_per_step_by_class={f"{'x'.join(map(str, k[0]))} {k[1]}": round(v * 1e3, 4) for k, v in per.items()})


def run_single_c2(cfa, steps, warmup):
    from channelflow_amd.fields import random_field
    nu = 1.0 / 2800.0
    Nx, Ny, Nz = 192, 129, 192
    Lx, Lz = 4 * np.pi, 2 * np.pi
    u0 = random_field(Nx, Ny, Nz, Lx, Lz, norm=0.2, decay=0.5, seed=2024)
    fl = poiseuille(nu, 0.005)
    dns = cfa.DNS([cfa.FlowField(Nx, Ny, Nz, 3, Lx, Lz, data=u0), cfa.FlowField(Nx, Ny, Nz, 1, Lx, Lz)],
                  cfa.DNSFlags(**fl))
    dns.advance_resident(warmup)
    dns.profile(True)
    t0 = time.perf_counter()
    dns.advance_resident(steps)
    wall = time.perf_counter() - t0
    kt = dns.kernel_times()
    dns.profile(False)
    kb = kernel_bytes((Nx, Ny, Nz), "Rotational")
    per = {}
    for k, (ms, cnt) in kt.items():
        if cnt > 0 and k in kb:
            per[k] = dict(ms_per_step=round(ms / steps, 4), GBs=round(kb[k] * cnt / (ms * 1e-3) / 1e9, 1))
    return dict(value=Nx * Ny * Nz * steps / wall, ms_per_step=wall / steps * 1e3, kernels=per,
                workload="plane Poiseuille Re_bulk=2800, 192x129x192, rotational NL, SBDF3 dt=0.005")


def host_info():
    """CPU model (/proc/cpuinfo, as lscpu prints it) and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(cpu_model=model, logical_cpus=os.cpu_count(), usable_cpus=len(os.sched_getaffinity(0)))


def _oracle_member_step_time(args):
    """Seconds per SBDF3 step of the oracle on one member: construction and the
    2 SMRK2 start steps untimed (dns.cpp:157-161), then at least `min_steps`
    steps and `min_s` seconds timed, stopping after `max_s` seconds."""
    mem, min_steps, min_s, max_s = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cfref as R
    Lx, Lz = mem.get("box", (LX_C, LZ_C))
    g = R.Grid(*mem["grid"], Lx, Lz)
    ref = R.OracleDNS(g, R.make_flags(**mem["flags"]), member_field(mem, Lx, Lz), g.zeros(1))
    ref.advance(2)
    t0 = time.perf_counter()
    n = 0
    while n < min_steps or (time.perf_counter() - t0) < min_s:
        ref.advance(1)
        n += 1
        if time.perf_counter() - t0 > max_s:
            break
    return mem["grid"], (time.perf_counter() - t0) / n, n


def cpu_baseline(members, steps_timed=2, budget_s=25.0, min_s=5.0, allcore_s=8.0, max_workers=16):
    """The oracle (oracle/cpu_ref.cpp, FP64, plain mixed-radix FFT) timed as
    tools/benchmark.cpp:74-146 times channelflow, in two legs:
      1 core: one member of each grid class in turn; ensemble time = sum over
        members of their class's time per step.
      all cores: one member per core, all at once (members spread over the
        ensemble so the classes keep its proportions), each timed as above;
        ensemble time = sum over members of their class's loaded time per step
        / cores (the members packed perfectly onto the cores).
    Runs before the GPU is touched, so that its worker processes fork from a
    process with no device state."""
    import multiprocessing as mp
    info = host_info()
    per = {}
    spent = 0.0
    for mem in members:
        if mem["grid"] in per:
            continue
        t0 = time.perf_counter()
        _, per[mem["grid"]], _ = _oracle_member_step_time((mem, steps_timed, min_s, budget_s))
        spent += time.perf_counter() - t0
    t_ens = sum(per[m["grid"]] for m in members)
    pts = sum(np.prod(m["grid"]) for m in members)
    workers = max(1, min(max_workers, info["usable_cpus"]))
    picks = [members[(k * len(members)) // workers] for k in range(workers)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_oracle_member_step_time, [(m, steps_timed, allcore_s, 3 * allcore_s) for m in picks])
    wall_all = time.perf_counter() - t0
    loaded = {}
    for g, t, _ in res:
        loaded.setdefault(g, []).append(t)
    loaded = {g: sum(v) / len(v) for g, v in loaded.items()}
    t_all = sum(loaded[m["grid"]] for m in members) / workers
    classes = ", ".join("x".join(map(str, k)) for k in per)
    return dict(value=pts / t_ens, unit="grid-pt*steps/s", cores=1, kind="port",
                sample=f"one member per grid class ({classes}): construction and 2 SMRK2 start steps untimed, "
                       f"then >= {steps_timed} SBDF3 steps and >= {min_s:.0f} s timed ({spent:.1f} s total) on 1 core; "
                       "ensemble time = sum of members' class times",
                ms_per_step_by_grid={"x".join(map(str, k)): round(v * 1e3, 2) for k, v in per.items()},
                all_cores={"value": pts / t_all, "cores": workers,
                           "sample": f"{workers} members at once, one per core ({len(picks)} picks spread over the "
                                     f"ensemble), each >= {allcore_s:.0f} s of SBDF3 steps after an untimed start "
                                     f"({wall_all:.1f} s wall); ensemble time = sum of members' loaded class times "
                                     "/ cores",
                           "ms_per_step_by_grid_loaded": {"x".join(map(str, k)): round(v * 1e3, 2)
                                                          for k, v in loaded.items()}},
                fft="the oracle's own mixed-radix complex FFT (no FFTW in this image)", **info)


def cpu_baseline_c2(min_steps=20):
    """The oracle on the 192x129x192 Poiseuille channel: construction and the 2
    SMRK2 start steps untimed, then >= min_steps SBDF3 steps timed and scaled
    linearly to 200 steps."""
    Nx, Ny, Nz = 192, 129, 192
    mem = dict(grid=(Nx, Ny, Nz), flags=poiseuille(1.0 / 2800.0, 0.005), norm=0.2, decay=0.5, seed=2024, member=0,
               box=(4 * np.pi, 2 * np.pi))
    t0 = time.perf_counter()
    _, per_step, n = _oracle_member_step_time((mem, min_steps, 0.0, 1e9))
    total = time.perf_counter() - t0
    return dict(value=Nx * Ny * Nz / per_step, unit="grid-pt*steps/s", cores=1, kind="port",
                seconds_per_step=per_step, seconds_for_200_steps_scaled=200 * per_step,
                sample=f"{n} SBDF3 steps timed after an untimed construction and SMRK2 start ({total:.0f} s in all), "
                       "scaled linearly to 200 steps", **host_info())


def roofline(members, kt, steps):
    """Dominant kernel of the batched run: algorithmic bytes of its launches
    over `steps` profiled steps of every grid class / their summed HIP-event
    durations (events recorded on the class's stream around every launch,
    one class at a time)."""
    totals = {}
    for mem in members:
        for k, b in kernel_bytes(mem["grid"], mem["flags"]["nonlinearity"]).items():
            totals[k] = totals.get(k, 0.0) + b * steps
    dom = max((k for k in kt if kt[k][1] > 0 and k in totals), key=lambda k: kt[k][0])
    ms, cnt = kt[dom]
    achieved = totals[dom] / (ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c3.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(dom)
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "launches": cnt,
            "avg_launch_ms": ms / cnt, "bytes_per_launch": totals[dom] / cnt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3", choices=["c3", "c4"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-c2", action="store_true",
                    help="also time the oracle on the 192x129x192 channel (>= 20 steps, about 2 minutes)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-ms", action="store_true", help="skip the multiple-shooting call (C4)")
    ap.add_argument("--c2-only", action="store_true",
                    help="run only the plain 192x129x192 channel and print its line (for profiling)")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--serve-budget", type=float, default=64e9, help="HBM budget of the server, bytes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain device copies of known size (for scaling HBM counters)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic_<workload>.json from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.traffic_from:
        cmd = ("rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -f csv -- python bench.py --workload "
               f"{args.workload} --steps 5 --warmup 3 --no-cpu --no-serve --no-c2 --no-plain --no-ms --copy-calibration")
        traffic_from_counters(*args.traffic_from, os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json"),
                              cmd)
        return
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(ensembles would shard by member across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    members = c3_members() if args.workload == "c3" else c4_members()
    cpu = None
    if not args.no_cpu and not args.c2_only:  # before the GPU is touched (its workers fork from this process)
        print("bench: cpu baseline (1 core, then all cores)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(members)
        if args.cpu_c2:
            print("bench: cpu baseline of the 192x129x192 channel", file=sys.stderr, flush=True)
            cpu["c2"] = cpu_baseline_c2()
        print(f"bench: cpu {json.dumps(cpu)}", file=sys.stderr, flush=True)

    import channelflow_amd as cfa
    if args.c2_only:
        print(json.dumps({"single_c2": run_single_c2(cfa, args.steps, args.warmup)}), flush=True)
        return
    label = ("C3 ensemble: 64 plane Couette Re=400 integrations (24x 32x33x32, 24x 48x49x48, 16x 64x65x64), "
             "skew-symmetric NL, bulk velocity, SBDF3 dt=0.02" if args.workload == "c3" else
             "C4 multi-shooting: 16 base fields x (1 + 8 perturbed copies), 144 plane Couette Re=400 "
             "integrations at 32x49x32, skew-symmetric NL, bulk velocity, SBDF3 dt=0.02")
    print(f"bench: {args.workload} batched, {len(members)} members, {args.steps} timed steps", file=sys.stderr,
          flush=True)
    prof_steps = max(10, min(50, args.steps))
    b = run_batched(cfa, members, args.steps, args.warmup, prof_steps)
    print(f"bench: batched {b['value']:.4g} grid-pt*steps/s, {b['wall'] / args.steps * 1e3:.3f} ms/step",
          file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "grid-pt*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded divergence-free perturbations)",
        "config": {"workload": label, "members": len(members), "grid_classes": b["classes"],
                   "grid_points_per_step": int(b["pts"]), "parallelism": "batched members, one GPU"},
        # members whose velocity stopped being finite (the reference's batch
        # client stops them too): removed, and not counted from the timed
        # region they failed in
        "members_stopped_nonfinite": b["dropped_nonfinite"],
        "ensemble_steps_run": b["steps_run"],
        "graphs": b["graphs"],
        "roofline": roofline(members, b["kt"], prof_steps),
        "kernels_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in b["kt"].items() if v[1] > 0},
        "kernel_profile_steps": prof_steps,
    }
    B_step = sum(16.0 * 25 * live(members[i]["grid"]) for i in b["timed_members"])
    out["step_roofline"] = {"bytes_per_step": B_step, "achieved": B_step / (b["wall"] / args.steps) / 1e9,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": B_step / (b["wall"] / args.steps) / 1e9 / HBM_PEAK_GBS}
    if not args.no_plain:
        print("bench: plain single-field path", file=sys.stderr, flush=True)
        p = run_plain(cfa, members, max(20, args.steps // 4), args.warmup)
        p["batched_over_plain"] = b["value"] / p["value"]
        print(f"bench: plain {p['value']:.4g} grid-pt*steps/s ({p['ms_per_step_by_grid']})", file=sys.stderr,
              flush=True)
        out["plain_gpu"] = p
    if not args.no_serve:
        print("bench: serving", file=sys.stderr, flush=True)
        out["served"] = run_served(cfa, args.serve_requests, args.serve_rate, 4242, args.serve_budget)
        print(f"bench: served {json.dumps(out['served'])}", file=sys.stderr, flush=True)
        print("bench: serving, every request queued at t=0", file=sys.stderr, flush=True)
        sat = run_served(cfa, args.serve_requests, None, 4242, args.serve_budget)
        sat["arrival_rate_per_s"] = "all at t=0"
        out["served"]["saturated"] = sat
        print(f"bench: served saturated {json.dumps(sat)}", file=sys.stderr, flush=True)
        # the same requests under a budget that binds: admission (with
        # backfilling) decides how many run at once
        print("bench: serving, every request queued at t=0, 4 GB budget", file=sys.stderr, flush=True)
        tight = run_served(cfa, args.serve_requests, None, 4242, 4e9)
        tight["arrival_rate_per_s"] = "all at t=0"
        out["served"]["budget_4GB"] = tight
        print(f"bench: served with a 4 GB budget {json.dumps(tight)}", file=sys.stderr, flush=True)
        if not args.no_plain:
            print("bench: served request mix on the plain path", file=sys.stderr, flush=True)
            sp = plain_for_requests(cfa, args.serve_requests, 4242, 20, 5)
            out["served"]["plain_gpu_same_requests"] = sp
            out["served"]["served_over_plain"] = out["served"]["value"] / sp["value"]
            sat["served_over_plain"] = sat["value"] / sp["value"]
    if not args.no_ms:
        print("bench: multiple-shooting call (C4)", file=sys.stderr, flush=True)
        out["multishooting"] = run_multishooting(cfa)
        print(f"bench: multishooting {json.dumps(out['multishooting'])}", file=sys.stderr, flush=True)
    if not args.no_c2:
        print("bench: single 192x129x192", file=sys.stderr, flush=True)
        out["single_c2"] = run_single_c2(cfa, 50, 10)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = {"one_core": b["value"] / cpu["value"], "all_cores": b["value"] / cpu["all_cores"]["value"]}
    if args.copy_calibration:
        copy_calibration()
        combine_calibration(cfa)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 counters
# ---------------------------------------------------------------------------
COPY_CAL_BYTES = 1 << 28


def copy_calibration(nbytes=COPY_CAL_BYTES, reps=3):
    """Device-to-device copies of nbytes by the HIP runtime's blit kernel. A
    profiled run reads its FETCH_SIZE / WRITE_SIZE counters against the known
    nbytes read and written per copy to scale the counters of the DNS kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]

    def chk(rc, what):
        if rc != 0:
            raise SystemExit(f"copy calibration: {what} failed ({rc})")

    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    chk(hip.hipMalloc(ctypes.byref(a), nbytes), "hipMalloc")
    chk(hip.hipMalloc(ctypes.byref(b), nbytes), "hipMalloc")
    chk(hip.hipMemset(a, 1, nbytes), "hipMemset")
    for _ in range(reps):
        chk(hip.hipMemcpy(b, a, nbytes, 3), "hipMemcpy")   # 3 = hipMemcpyDeviceToDevice
    chk(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    chk(hip.hipFree(a), "hipFree")
    chk(hip.hipFree(b), "hipFree")
    print(f"copy calibration: {reps} device copies of {nbytes} bytes", file=sys.stderr, flush=True)


COMBINE_GRID = (192, 129, 192)


def combine_field_doubles(grid):
    """Doubles in one 3-component field of the compact live-mode layout."""
    Nx, Ny, Nz = grid
    return 2 * 3 * Ny * (2 * (Nx // 3 - 1) + 1) * (Nz // 3)


def combine_calibration(cfa, steps=2):
    """A CNRK2 integration on COMBINE_GRID: every Runge-Kutta substep runs one
    k_combine of two terms over one 3-component field (dnsalgo.cpp:420-446),
    which reads exactly 2 and writes exactly 1 such field. A counter run checks
    the copy-based scale of FETCH_SIZE / WRITE_SIZE on it."""
    from channelflow_amd.fields import random_field
    Nx, Ny, Nz = COMBINE_GRID
    Lx, Lz = 4 * np.pi, 2 * np.pi
    fl = dict(poiseuille(1.0 / 2800.0, 0.005), timestepping="CNRK2", initstepping="CNRK2")
    u0 = random_field(Nx, Ny, Nz, Lx, Lz, norm=0.2, decay=0.5, seed=2024)
    dns = cfa.DNS([cfa.FlowField(Nx, Ny, Nz, 3, Lx, Lz, data=u0), cfa.FlowField(Nx, Ny, Nz, 1, Lx, Lz)],
                  cfa.DNSFlags(**fl))
    dns.advance_resident(steps)
    dns.state()
    print(f"combine calibration: {3 * steps} k_combine launches over {COMBINE_GRID}", file=sys.stderr, flush=True)


# device kernels that fill a step slot of another name (wavefront-local FFT passes)
KERNEL_ALIAS = {"k_zwave": "k_zpass", "k_xwinv": "k_xinv", "k_xwfwd": "k_xfwd"}


def _counter_rows(path, counter):
    """(short kernel name, value) per dispatch from a rocprofv3 counter_collection.csv."""
    import csv
    import re
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"]
            m = re.search(r"cfg::(k_\w+?)(?:_wave|_w|_t)?(?:<|\()", name)  # variants count as their kernel
            short = m.group(1) if m else name.split("(")[0].strip()
            rows.append((KERNEL_ALIAS.get(short, short), float(r["Counter_Value"])))
    return rows


def traffic_from_counters(fetch_csv, write_csv, out_path, command):
    """HBM bytes per launch of each DNS kernel from one FETCH_SIZE and one
    WRITE_SIZE run (KiB per dispatch), each scaled by the same counter's reading
    on the plain copies of copy_calibration()."""
    per = {}
    scale = {}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        rows = _counter_rows(path, counter)
        # the calibration copies, not the small state uploads: FETCH_SIZE reads
        # about half a copy's bytes, so the cut is at a quarter
        cal = [v for n, v in rows if "copyBuffer" in n and v * 1024 > COPY_CAL_BYTES / 4]
        if not cal:
            raise SystemExit(f"{path}: no calibration copy dispatches of {COPY_CAL_BYTES} bytes")
        scale[key] = COPY_CAL_BYTES / (1024.0 * sum(cal) / len(cal))
        acc = {}
        for n, v in rows:
            if n.startswith("k_"):
                acc.setdefault(n, []).append(v)
        for n, vs in acc.items():
            per.setdefault(n, {})[key] = 1024.0 * scale[key] * sum(vs) / len(vs)
    out = {n: d.get("fetch", 0.0) + d.get("write", 0.0) for n, d in per.items()}
    if "k_combine" in per:  # exact bytes: 2 fields read, 1 written per launch (combine_calibration)
        nf = 8.0 * combine_field_doubles(COMBINE_GRID)
        cf = per["k_combine"]
        out["_combine_check"] = {"grid": COMBINE_GRID, "read_counter": cf.get("fetch"), "read_exact": 2 * nf,
                                 "read_ratio": cf.get("fetch", 0.0) / (2 * nf), "write_counter": cf.get("write"),
                                 "write_exact": nf, "write_ratio": cf.get("write", 0.0) / nf}
    out["_per_kernel"] = per
    out["_scale"] = scale
    out["_command"] = command
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if not k.startswith("_")}))


if __name__ == "__main__":
    main()
