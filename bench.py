"""Benchmark of the MI355X DNS time step (one GPU).

Workload (default): plane Poiseuille flow, Re_bulk ~ 2800 (Re_tau ~ 180 class),
box Lx = 4 pi, Lz = 2 pi, grid 192 x 129 x 192, FP64, rotational nonlinearity,
2/3 dealiasing, constant mean pressure gradient, SBDF3 with dt = 0.005 and the
SMRK2 start, initial field = parabolic base flow + seeded divergence-free
perturbation of L2 norm 0.2 (spectral decay 0.5). nu = 1/2800 and
dPdx = -3 nu, the gradient whose laminar profile has bulk velocity 1
(Re_bulk = 2800; centreline 1.5). The gradient -(180 nu)^2 of a turbulent
Re_tau = 180 channel would drive a laminar centreline of 5.8 and put dt = 0.005
past the explicit-term stability limit (CFL ~ 0.9).

A step is one DNS::advance step of that single field; the velocity, pressure
and multistep history are resident in HBM before the timed region starts.
Prints one JSON line (metric, value in grid-point*steps/s, roofline of the
dominant kernel from HIP-event timings of the same timed region, and the
oracle's CPU rate on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def workload(name):
    if name == "c2":
        nu = 1.0 / 2800.0
        return dict(
            name="c2", label="plane Poiseuille Re_bulk=2800, 192x129x192, rotational NL, SBDF3 dt=0.005",
            grid=(192, 129, 192), Lx=4 * np.pi, Lz=2 * np.pi, norm=0.2, decay=0.5, seed=2024,
            flags=dict(baseflow="LaminarBase", constraint="PressureGradient", dPdx=-3.0 * nu, nu=nu,
                       timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational",
                       dealiasing="DealiasXZ", taucorrection=True, dt=0.005))
    if name == "c1":
        return dict(
            name="c1", label="plane Couette Re=400, 32x33x32, skew-symmetric NL, bulk velocity, SBDF3 dt=0.02",
            grid=(32, 33, 32), Lx=2 * np.pi / 1.14, Lz=2 * np.pi / 2.5, norm=0.1, decay=0.3, seed=2024,
            flags=dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity",
                       Ubulk=0.0, nu=1.0 / 400, timestepping="SBDF3", initstepping="SMRK2",
                       nonlinearity="SkewSymmetric", dealiasing="DealiasXZ", taucorrection=True, dt=0.02))
    raise SystemExit(f"unknown workload {name}")


def kernel_bytes(wl, method):
    """Compulsory HBM bytes per launch of each kernel (inputs read once, outputs written once)."""
    Nx, Ny, Nz = wl["grid"]
    kxd, kzd = Nx // 3 - 1, Nz // 3 - 1
    M = (2 * kxd + 1) * (kzd + 1)
    L = 16.0 * Ny * M                     # one complex component array in the compact spectral layout
    X = 16.0 * Ny * Nx * (kzd + 1)        # one complex field, x physical, kz spectral
    nyout = {"Rotational": 6, "SkewSymmetric": 6, "Convection": 6, "Divergence": 3}[method]
    nxout = {"Rotational": 6, "SkewSymmetric": 9, "Convection": 9, "Divergence": 3}[method]
    nzout = {"Rotational": 3, "SkewSymmetric": 9, "Convection": 3, "Divergence": 6}[method]
    return {
        "k_yinv": (3 + nyout) * L,
        "k_xinv": nyout * L + nxout * X,
        "k_zpass": nxout * X + nzout * X,
        "k_xfwd": nzout * X + nzout * L,
        "k_yfwd": nzout * L + 3 * L,
        "k_combine": (2 * 3 * 3 + 3) * L,  # SBDF3: u_n..u_{n-2}, f_n..f_{n-2} in, rhs out
        "k_tau_solve": (3 + 3 + 1) * L,     # rhs in, u and q out
    }, L / 16.0


def cpu_baseline(wl, steps_timed, budget_s=30.0):
    """Oracle (oracle/cpu_ref.cpp, single thread) on the same workload: untimed
    construction and SMRK2 start (dns.cpp:157-161), then a timed block of SBDF3
    steps, as tools/benchmark.cpp:74-146 times channelflow."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cfref as R
    from channelflow_amd.fields import random_field
    Nx, Ny, Nz = wl["grid"]
    g = R.Grid(Nx, Ny, Nz, wl["Lx"], wl["Lz"])
    u0 = random_field(Nx, Ny, Nz, wl["Lx"], wl["Lz"], norm=wl["norm"], decay=wl["decay"], seed=wl["seed"])
    dns = R.OracleDNS(g, R.make_flags(**wl["flags"]), u0, g.zeros(1))
    dns.advance(2)  # SMRK2 start, untimed
    t0 = time.perf_counter()
    n = 0
    while n < steps_timed:
        dns.advance(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=Nx * Ny * Nz * n / dt, unit="grid-pt*steps/s", cores=1, kind="port",
                sample=f"{wl['name']} grid {Nx}x{Ny}x{Nz}: construction and 2 SMRK2 start steps untimed, "
                       f"then {n} SBDF3 steps timed ({dt:.1f} s) on 1 core")


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
    print(f"copy calibration: {reps} device copies of {nbytes} bytes", flush=True)


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
            m = re.search(r"cfg::(k_\w+?)(?:_wave)?(?:<|\()", name)
            rows.append((m.group(1) if m else name.split("(")[0].strip(), float(r["Counter_Value"])))
    return rows


def traffic_from_counters(fetch_csv, write_csv, out_path, command):
    """HBM bytes per launch of each DNS kernel from one FETCH_SIZE and one
    WRITE_SIZE run (KiB per dispatch), each scaled by the same counter's reading
    on the plain copies of copy_calibration()."""
    per = {}
    scale = {}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        rows = _counter_rows(path, counter)
        cal = [v for n, v in rows if "copyBuffer" in n and v * 1024 > COPY_CAL_BYTES / 2]
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
    out["_per_kernel"] = per
    out["_scale"] = scale
    out["_command"] = command
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if not k.startswith("_")}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain device copies of known size (for scaling HBM counters)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic_<workload>.json from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.traffic_from:
        cmd = ("rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -f csv -- python bench.py --workload "
               f"{args.workload} --steps 5 --warmup 3 --no-cpu --copy-calibration")
        traffic_from_counters(*args.traffic_from, os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json"),
                              cmd)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(ensembles shard by member across GPUs; see the sharding notes of the design)", file=sys.stderr)
        sys.exit(2)

    import channelflow_amd as cfa
    from channelflow_amd.fields import random_field
    wl = workload(args.workload)
    Nx, Ny, Nz = wl["grid"]
    u0 = random_field(Nx, Ny, Nz, wl["Lx"], wl["Lz"], norm=wl["norm"], decay=wl["decay"], seed=wl["seed"])
    flags = cfa.DNSFlags(**wl["flags"])
    fu = cfa.FlowField(Nx, Ny, Nz, 3, wl["Lx"], wl["Lz"], data=u0)
    fq = cfa.FlowField(Nx, Ny, Nz, 1, wl["Lx"], wl["Lz"])
    dns = cfa.DNS([fu, fq], flags)
    dns.advance_resident(args.warmup)       # SMRK2 start + warm-up SBDF3 steps, untimed

    dns.profile(True)                       # HIP events around every launch on the DNS's stream
    t0 = time.perf_counter()
    dns.advance_resident(args.steps)        # returns after the stream has drained
    t1 = time.perf_counter()
    kt = dns.kernel_times()
    dns.profile(False)
    wall = t1 - t0
    u, _ = dns.state()
    if not np.all(np.isfinite(u)):
        raise SystemExit("non-finite state after the timed steps")

    pts = Nx * Ny * Nz
    value = pts * args.steps / wall
    kb, Lc = kernel_bytes(wl, wl["flags"]["nonlinearity"])
    dom = max((k for k in kt if kt[k][1] > 0), key=lambda k: kt[k][0])
    avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    achieved = kb[dom] / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{wl['name']}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(dom)
    B_step = 16.0 * 25 * Lc                 # compulsory state bytes per SBDF3 step
    out = {
        "metric": "grid-pt·steps/s batched+served vs plain GPU at 1e-10 rel-L2; p50/p99 latency",
        "value": value,
        "unit": "grid-pt*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded divergence-free perturbation)",
        "config": {"workload": wl["label"], "grid": list(wl["grid"]), "steps_per_call": args.steps,
                   "parallelism": "single field, one GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "step_roofline": {"bytes_per_step": B_step, "achieved": B_step / (wall / args.steps) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": B_step / (wall / args.steps) / 1e9 / HBM_PEAK_GBS},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1] > 0},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps)
    if args.copy_calibration:
        copy_calibration()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
