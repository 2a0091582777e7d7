"""GPU parity: the HIP path through the C ABI against the oracle (oracle/cpu_ref.cpp).

Tolerances, the bar the FP64 path is held to: relative L2 <= 1e-12 per velocity component for one
nonlinear-term or one tau-solve call; <= 1e-10 after 20 time steps; the
reference's own golden trajectory (tests/timeIntegrationTest.cpp) at its own
bar, L2Dist <= 1e-13.
"""
import json
import os

import numpy as np
import pytest

import cfref as R
import channelflow_amd as cfa
from channelflow_amd.fields import random_field

HERE = os.path.dirname(os.path.abspath(__file__))
LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5

pytestmark = pytest.mark.gpu


def rel_l2_per_component(g, a, b):
    """L2Dist(a_i, b_i) / L2Norm(b_i) per component (diffops.cpp:353-487 norms)."""
    out = []
    for i in range(a.shape[0]):
        d = np.sqrt(R.l2dist2(g, a[i:i + 1], b[i:i + 1]))
        n = np.sqrt(R.l2norm2(g, b[i:i + 1]))
        out.append(d / n if n > 0 else d)
    return np.array(out)


def live_mask(Nx, Nz, kxd, kzd):
    kx = np.fft.fftfreq(Nx, 1.0 / Nx).astype(int)
    m = np.zeros((Nx, Nz // 2 + 1), dtype=bool)
    m[np.abs(kx) <= kxd, :kzd + 1] = True
    return m


@pytest.mark.parametrize("method", ["Rotational", "Convection", "Divergence", "SkewSymmetric"])
@pytest.mark.parametrize("dealias", ["DealiasXZ", "NoDealiasing"])
@pytest.mark.parametrize("grid", [(24, 17, 18), (16, 17, 24), (32, 17, 48), (16, 19, 192), (8, 9, 16), (12, 11, 32),
                                  (12, 13, 64), (8, 9, 96), (8, 9, 128), (4, 7, 256), (48, 9, 16), (64, 9, 12),
                                  (96, 9, 8), (128, 7, 16), (192, 7, 8), (256, 7, 8)])
def test_nonlinear_single_call(gpu, method, dealias, grid):
    """z-pass: Nz = 18 takes the runtime-length LDS FFTs, 24 the compile-time
    LDS ones, the others the wavefront FFTs (Nx = 12 at Nz = 32 and 64, and
    Nx = 8 at Nz = 16, leave the last wavefront of each y partly without
    lines). x-FFTs: Nx = 16, 32, 48, 64, 96, 128, 192, 256 take the wavefront
    FFTs (lines are kz, so most of these end in a partial wavefront), the
    others the LDS ones."""
    Nx, Ny, Nz = grid
    g = R.Grid(Nx, Ny, Nz, LX_C, LZ_C)
    u = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.3, decay=0.5, seed=11, dealias=(dealias == "DealiasXZ"))
    fl = dict(nonlinearity=method, dealiasing=dealias, baseflow="LaminarBase", uupperwall=1.0, ulowerwall=-1.0,
              constraint="BulkVelocity", nu=1 / 400, Vsuck=0.0)
    f_ref, _ = R.nonlinear(g, R.make_flags(**fl), u)
    f_gpu = cfa.nonlinear(cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u), cfa.DNSFlags(**fl))
    kxd = Nx // 3 - 1 if dealias == "DealiasXZ" else Nx // 2 - 1
    kzd = Nz // 3 - 1 if dealias == "DealiasXZ" else Nz // 2 - 1
    m = live_mask(Nx, Nz, kxd, kzd)
    fr = np.where(m[None, None], f_ref, 0)
    err = rel_l2_per_component(g, f_gpu, fr)
    print(grid, method, dealias, "rel L2 per component", err, flush=True)
    assert np.all(err <= 1e-12), err


@pytest.mark.parametrize("method", ["Rotational", "Convection", "Divergence", "SkewSymmetric"])
@pytest.mark.parametrize("Ny", [33, 49, 65, 97, 129, 161, 193, 257, 289])
def test_nonlinear_chebyshev_sizes(gpu, method, Ny):
    """y transforms at every row-tile count of the matrix-core kernels (16
    upper points or coefficient pairs per tile: Ny = 33 one tile, 49 and 65
    two, 97 three, 129 four, 161 and 193 six, 257 eight), including the
    chunk-wise LDS staging of the forward transform's divergence forms at six
    and eight tiles, and the vector-tile kernels above 257 (Ny = 289). A
    single nonlinear-term call against the oracle."""
    Nx, Nz = 8, 8
    g = R.Grid(Nx, Ny, Nz, LX_C, LZ_C)
    u = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.3, decay=0.5, seed=13, dealias=True)
    fl = dict(nonlinearity=method, dealiasing="DealiasXZ", baseflow="LaminarBase", uupperwall=1.0,
              ulowerwall=-1.0, constraint="BulkVelocity", nu=1 / 400, Vsuck=0.0)
    f_ref, _ = R.nonlinear(g, R.make_flags(**fl), u)
    f_gpu = cfa.nonlinear(cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u), cfa.DNSFlags(**fl))
    m = live_mask(Nx, Nz, Nx // 3 - 1, Nz // 3 - 1)
    fr = np.where(m[None, None], f_ref, 0)
    err = rel_l2_per_component(g, f_gpu, fr)
    print(Ny, method, "rel L2 per component", err, flush=True)
    assert np.all(err <= 1e-12), err


@pytest.mark.parametrize("constraint", ["PressureGradient", "BulkVelocity"])
@pytest.mark.parametrize("taucorr", [True, False])
@pytest.mark.parametrize("grid", [(16, 33, 16), (8, 25, 8), (8, 65, 8), (8, 97, 8), (8, 129, 8), (6, 193, 6),
                                  (6, 257, 6), (6, 321, 6)])
def test_solve_single_call(gpu, constraint, taucorr, grid):
    """Ny = 25, 33, 65 take 16-lane segments with 1, 2, 3 even coefficients per
    lane; 97, 193, 257 64-lane segments with 1, 2, 3; 129 32-lane segments
    with 3. Ny = 321 needs more than 64 KB of LDS per block."""
    Nx, Ny, Nz = grid
    g = R.Grid(Nx, Ny, Nz, LX_C, LZ_C)
    rng = np.random.default_rng(5)
    rhs = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=1.0, decay=0.7, seed=3) + \
        1e-3 * (rng.standard_normal((3, Ny, Nx, Nz // 2 + 1)) + 1j * rng.standard_normal((3, Ny, Nx, Nz // 2 + 1)))
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.1, decay=0.5, seed=4)
    fl = dict(constraint=constraint, taucorrection=taucorr, nu=1 / 400, dPdx=-0.01, Ubulk=0.6,
              baseflow="LaminarBase")
    lam = (11.0 / 6.0) / 0.02
    u_ref, q_ref, dp_ref = R.solve(g, R.make_flags(**fl), lam, u0, rhs)
    u_gpu, q_gpu, dp_gpu = cfa.solve(cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u0), rhs, cfa.DNSFlags(**fl), lam)
    eu = rel_l2_per_component(g, u_gpu, u_ref)
    eq = rel_l2_per_component(g, q_gpu, q_ref)
    print(constraint, taucorr, "u", eu, "q", eq, "dPdx", dp_gpu, dp_ref, flush=True)
    assert np.all(eu <= 1e-12) and np.all(eq <= 1e-12)
    if constraint == "BulkVelocity":
        assert abs(dp_gpu - dp_ref) <= 1e-12 * max(1.0, abs(dp_ref))


C1_FLAGS = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
                timestepping="SBDF3", initstepping="SMRK2", nonlinearity="SkewSymmetric", dealiasing="DealiasXZ",
                taucorrection=True, nu=1 / 400, dt=0.02)


def run_pair(Nx, Ny, Nz, Lx, Lz, flags, u0, steps, chunks=1):
    g = R.Grid(Nx, Ny, Nz, Lx, Lz)
    ref = R.OracleDNS(g, R.make_flags(**flags), u0, g.zeros(1))
    fu = cfa.FlowField(Nx, Ny, Nz, 3, Lx, Lz, data=u0)
    fq = cfa.FlowField(Nx, Ny, Nz, 1, Lx, Lz)
    dns = cfa.DNS([fu, fq], cfa.DNSFlags(**flags))
    for _ in range(chunks):
        ref.advance(steps // chunks)
        dns.advance_resident(steps // chunks)
    ur, qr = ref.state()
    ug, qg = dns.state()
    return g, ur, qr, ug, qg, ref, dns


def test_dns_c1_20_steps(gpu):
    """C1 physics (Couette Re=400, skew-symmetric, bulk velocity, SBDF3 + SMRK2) at 32x33x32, 20 steps."""
    u0 = random_field(32, 33, 32, LX_C, LZ_C, norm=0.1, decay=0.3, seed=1)
    g, ur, qr, ug, qg, ref, dns = run_pair(32, 33, 32, LX_C, LZ_C, C1_FLAGS, u0, 20, chunks=2)
    eu = rel_l2_per_component(g, ug, ur)
    print("C1 20 steps rel L2 per component", eu, "t", dns.time(), ref.time, flush=True)
    assert np.all(eu <= 1e-10), eu
    assert abs(dns.time() - ref.time) < 1e-12
    assert abs(dns.dPdx() - ref.dPdx) <= 1e-10 * max(1.0, abs(ref.dPdx))


@pytest.mark.parametrize("method,stepping,init", [("Rotational", "SBDF2", "SMRK2"), ("Convection", "SBDF4", "SMRK2"),
                                                  ("Divergence", "SBDF1", "SMRK2"), ("Alternating", "CNAB2", "SMRK2"),
                                                  ("Rotational", "CNRK2", "CNRK2"), ("SkewSymmetric", "SMRK2", "SMRK2"),
                                                  ("SkewSymmetric", "CNFE1", "SMRK2")])
def test_dns_schemes(gpu, method, stepping, init):
    flags = dict(C1_FLAGS, nonlinearity=method, timestepping=stepping, initstepping=init)
    u0 = random_field(24, 25, 24, LX_C, LZ_C, norm=0.1, decay=0.3, seed=2)
    g, ur, qr, ug, qg, ref, dns = run_pair(24, 25, 24, LX_C, LZ_C, flags, u0, 12, chunks=3)
    eu = rel_l2_per_component(g, ug, ur)
    print(method, stepping, init, "rel L2", eu, flush=True)
    assert np.all(eu <= 1e-10), eu


CFL_CASES = {
    "couette_skew": (C1_FLAGS, (24, 25, 24)),
    "poiseuille_rot_nodealias": (dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 / 2000,
                                      nu=1 / 2000, dt=0.01, timestepping="SBDF3", initstepping="SMRK2",
                                      nonlinearity="Rotational", dealiasing="NoDealiasing", taucorrection=True),
                                 (16, 17, 16)),
    "suction_rot": (dict(baseflow="SuctionBase", constraint="PressureGradient", dPdx=0.0, nu=0.0025, Vsuck=0.0025,
                         dt=0.025, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational",
                         dealiasing="DealiasXZ", taucorrection=True), (24, 19, 24)),
}


@pytest.mark.parametrize("case", sorted(CFL_CASES))
def test_cfl_matches_oracle(gpu, case):
    """DNS::CFL (dnsalgo.cpp:63-67): the device's value on the same initial state <= 1e-12
    relative to the oracle's, and again after 6 steps of each."""
    flags, (Nx, Ny, Nz) = CFL_CASES[case]
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.2, decay=0.4, seed=31, dealias=flags["dealiasing"] != "NoDealiasing")
    g, ur, qr, ug, qg, ref, dns = run_pair(Nx, Ny, Nz, LX_C, LZ_C, flags, u0, 0)
    c0_ref, c0_gpu = ref.cfl(), dns.CFL()
    ref.advance(6)
    dns.advance_resident(6)
    c6_ref, c6_gpu = ref.cfl(), dns.CFL()
    print(f"{case}: CFL t=0 oracle {c0_ref:.15e} gpu {c0_gpu:.15e}; t=6dt oracle {c6_ref:.15e} gpu {c6_gpu:.15e}",
          flush=True)
    assert abs(c0_gpu - c0_ref) <= 1e-12 * abs(c0_ref)
    assert abs(c6_gpu - c6_ref) <= 1e-12 * abs(c6_ref)


def test_variable_dt_loop(gpu):
    """simulateflow's loop (programs/simulateflow.cpp:131-142): advance n steps, CFL,
    TimeStep.adjust, reset_dt on a change -- same dt sequence and states within
    1e-10 of the oracle driven the same way."""
    Nx, Ny, Nz = 24, 25, 24
    flags = dict(C1_FLAGS, dt=0.05)
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.1, decay=0.3, seed=41)
    g, _, _, _, _, ref, dns = run_pair(Nx, Ny, Nz, LX_C, LZ_C, flags, u0, 0)
    ts_g = cfa.TimeStep(0.05, 0.001, 0.2, 0.5, 0.1, 0.2)
    ts_r = cfa.TimeStep(0.05, 0.001, 0.2, 0.5, 0.1, 0.2)
    changes = 0
    for k in range(4):
        dns.advance_resident(ts_g.n())
        ref.advance(ts_r.n())
        cg, cr = dns.CFL(), ref.cfl()
        assert abs(cg - cr) <= 1e-10 * abs(cr)
        ag, ar = ts_g.adjust(cg), ts_r.adjust(cr)
        assert ag == ar and ts_g.n() == ts_r.n()
        if ag:
            dns.reset_dt(ts_g.dt())
            ref.reset_dt(ts_r.dt())
            changes += 1
        print(f"interval {k}: CFL {cg:.6f} n {ts_g.n()} dt {ts_g.dt():.6f} t {dns.time():.6f}", flush=True)
    assert changes >= 1
    eu = rel_l2_per_component(g, dns.state()[0], ref.state()[0])
    print("variable dt: rel L2 per component", eu, flush=True)
    assert np.all(eu <= 1e-10), eu
    assert abs(dns.time() - ref.time) < 1e-12


PRESSURE_CASES = {
    "couette_conv": dict(C1_FLAGS, nonlinearity="Convection"),
    "couette_skew": C1_FLAGS,
    "poiseuille_rot": CFL_CASES["poiseuille_rot_nodealias"][0] | dict(dealiasing="DealiasXZ"),
    "suction_rot": CFL_CASES["suction_rot"][0],
}


@pytest.mark.parametrize("case", sorted(PRESSURE_CASES))
def test_pressure_matches_oracle(gpu, case):
    """PressureSolver(u, Ubase, Wbase, nu, Vsuck, nonlinearity).solve(q, u): the device's
    q on the live modes <= 1e-12 relative to the oracle's, with the base flow taken
    from the flags and passed explicitly."""
    flags = PRESSURE_CASES[case]
    Nx, Ny, Nz = 24, 25, 24
    g = R.Grid(Nx, Ny, Nz, LX_C, LZ_C)
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.2, decay=0.4, seed=51)
    q_ref = R.pressure(g, R.make_flags(**flags), u0)
    fu = cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u0)
    fq = cfa.FlowField(Nx, Ny, Nz, 1, LX_C, LZ_C)
    cfa.PressureSolver(fu, flags=cfa.DNSFlags(**flags)).solve(fq, fu)
    ref = R.OracleDNS(g, R.make_flags(**flags), u0, g.zeros(1))
    U, W = ref.base()
    fq2 = cfa.FlowField(Nx, Ny, Nz, 1, LX_C, LZ_C)
    cfa.PressureSolver(fu, U, W, flags["nu"], flags.get("Vsuck", 0.0), flags["nonlinearity"]).solve(fq2, fu)
    mask = live_mask(Nx, Nz, Nx // 3 - 1, Nz // 3 - 1)
    a, b = fq.data[0][:, mask], q_ref[0][:, mask]
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    err2 = np.linalg.norm(fq2.data[0][:, mask] - b) / np.linalg.norm(b)
    print(f"{case}: pressure rel L2 vs oracle {err:.3e} (explicit base {err2:.3e}), |q| {np.linalg.norm(b):.3e}",
          flush=True)
    assert err <= 1e-12 and err2 <= 1e-12
    assert np.all(fq.data[0][:, ~mask] == 0)


@pytest.mark.parametrize("case", ["couette_skew", "poiseuille_rot"])
@pytest.mark.parametrize("Ny", [49, 97, 193, 257])
def test_pressure_chebyshev_sizes(gpu, case, Ny):
    """The pressure kernel's other segment shapes (Ny = 25 above takes 16-lane
    segments): 32-lane at Ny = 49, 64-lane with 1, 2, 3 coefficients per lane
    at 97, 193, 257; Ny = 257 needs more than 64 KB of LDS per block.

    The Poisson problem for q loses accuracy about as Ny^4: the oracle's own q
    moves by 5e-12 (Couette, skew-symmetric, Ny = 97) to 2e-10 (Ny = 257)
    when its input changes by 1e-15 relative. So the bar is 1e-12 or ten times
    the oracle's change under two such perturbations, whichever is larger."""
    flags = PRESSURE_CASES[case]
    Nx, Nz = 8, 8
    g = R.Grid(Nx, Ny, Nz, LX_C, LZ_C)
    u0 = random_field(Nx, Ny, Nz, LX_C, LZ_C, norm=0.2, decay=0.4, seed=52)
    fl = R.make_flags(**flags)
    q_ref = R.pressure(g, fl, u0)
    fu = cfa.FlowField(Nx, Ny, Nz, 3, LX_C, LZ_C, data=u0)
    fq = cfa.FlowField(Nx, Ny, Nz, 1, LX_C, LZ_C)
    cfa.PressureSolver(fu, flags=cfa.DNSFlags(**flags)).solve(fq, fu)
    mask = live_mask(Nx, Nz, Nx // 3 - 1, Nz // 3 - 1)
    a, b = fq.data[0][:, mask], q_ref[0][:, mask]
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    rng = np.random.default_rng(53)
    sens = 0.0
    for _ in range(2):
        qp = R.pressure(g, fl, u0 * (1 + 1e-15 * rng.standard_normal(u0.shape)))[0][:, mask]
        sens = max(sens, np.linalg.norm(qp - b) / np.linalg.norm(b))
    bar = max(1e-12, 10 * sens)
    print(f"{case} Ny={Ny}: pressure rel L2 vs oracle {err:.3e}, oracle's change under 1e-15 input noise "
          f"{sens:.3e}, bar {bar:.3e}", flush=True)
    assert err <= bar


def test_golden_trajectory(gpu):
    """timeIntegrationTest: uinit -> 440 SBDF3 steps -> ufinal, L2Dist <= 1e-13 (the reference's own bar)."""
    man = json.load(open(os.path.join(HERE, "golden", "manifest.json")))
    mi = man["uinit"]
    g = R.Grid(mi["Nx"], mi["Ny"], mi["Nz"], mi["Lx"], mi["Lz"], mi["a"], mi["b"])
    shape = tuple(mi["shape"])
    u0 = R.field_from_io(g, np.fromfile(os.path.join(HERE, "golden", "uinit.f64")).reshape(shape))
    uf = R.field_from_io(g, np.fromfile(os.path.join(HERE, "golden", "ufinal.f64")).reshape(shape))
    t = man["timeIntegrationTest"]
    flags = cfa.DNSFlags(baseflow="SuctionBase", timestepping="SBDF3", initstepping="SMRK2",
                         nonlinearity="Rotational", dealiasing="DealiasXZ", taucorrection=True,
                         constraint="PressureGradient", dPdx=0.0, dt=t["dt"], nu=t["nu"], Vsuck=t["Vsuck"])
    fu = cfa.FlowField(mi["Nx"], mi["Ny"], mi["Nz"], 3, mi["Lx"], mi["Lz"], data=u0)
    fq = cfa.FlowField(mi["Nx"], mi["Ny"], mi["Nz"], 1, mi["Lx"], mi["Lz"])
    dns = cfa.DNS([fu, fq], flags)
    for _ in range(11):  # T = 10 in unit steps of 40 (timeIntegrationTest.cpp:104-137)
        dns.advance([fu, fq], 40)
    d = np.sqrt(R.l2dist2(g, fu.data, uf))
    print("golden L2Dist", d, flush=True)
    assert d <= t["tolerance_L2Dist"], d


# 192x129x192 plane Poiseuille, Re_bulk = 2800 (nu = 1/2800, dPdx = -3 nu: laminar
# bulk velocity 1), rotational, 2/3 dealiasing, SBDF3 dt = 0.005 with the SMRK2 start
C2_WORKLOAD = dict(grid=(192, 129, 192), Lx=4 * np.pi, Lz=2 * np.pi,
                   flags=dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 / 2800, nu=1.0 / 2800,
                              timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational",
                              dealiasing="DealiasXZ", taucorrection=True, dt=0.005))


def test_c2_full_size_20_steps(gpu):
    """192x129x192 Poiseuille (bench workload), 20 steps, against the oracle's committed
    fixture (tests/golden/make_c2_reference.py): per-component norms and every
    coefficient of 12 sampled modes at 1e-10; divergence-free and no-slip at full size."""
    from channelflow_amd.fields import divergence
    meta = json.load(open(os.path.join(HERE, "golden", "c2_20steps.json")))
    ref = np.fromfile(os.path.join(HERE, "golden", "c2_20steps_modes.c128"), dtype=np.complex128)
    ref = ref.reshape(meta["shape"])
    wl = C2_WORKLOAD
    Nx, Ny, Nz = wl["grid"]
    u0 = random_field(Nx, Ny, Nz, wl["Lx"], wl["Lz"], norm=meta["norm"], decay=meta["decay"], seed=meta["seed"])
    fu = cfa.FlowField(Nx, Ny, Nz, 3, wl["Lx"], wl["Lz"], data=u0)
    fq = cfa.FlowField(Nx, Ny, Nz, 1, wl["Lx"], wl["Lz"])
    dns = cfa.DNS([fu, fq], cfa.DNSFlags(**wl["flags"]))
    dns.advance_resident(meta["steps"])
    u, q = dns.state()
    g = R.Grid(Nx, Ny, Nz, wl["Lx"], wl["Lz"])
    norms = np.array([np.sqrt(R.l2norm2(g, u[i:i + 1])) for i in range(3)])
    en = np.abs(norms - np.array(meta["l2norm_u"])) / np.array(meta["l2norm_u"])
    got = np.stack([u[:, :, kx % Nx, kz] for kx, kz in meta["modes"]], axis=0)
    em = np.array([np.linalg.norm(got[:, i] - ref[:, i]) / np.linalg.norm(ref[:, i]) for i in range(3)])
    div = np.abs(divergence(u, wl["Lx"], wl["Lz"])).max() / np.abs(u).max()
    walls = max(np.abs(u.sum(axis=1)).max(), np.abs((u * (-1.0) ** np.arange(Ny)[None, :, None, None]).sum(axis=1)).max())
    print("C2 20 steps: norm rel diff", en, "sampled-mode rel L2", em, "div", div, "wall", walls, flush=True)
    assert np.all(en <= 1e-10) and np.all(em <= 1e-10)
    assert div < 1e-11 and walls < 1e-12
    assert abs(dns.time() - meta["time"]) < 1e-12
