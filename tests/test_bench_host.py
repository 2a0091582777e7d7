"""Host logic of bench.py, no GPU: the one-GPU refusal, the live-mode count
behind the algorithmic bytes, the serving request stream, and the conversion
of rocprofv3 counters into HBM bytes per launch."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("grid", [(32, 33, 32), (48, 49, 48), (64, 65, 64), (192, 129, 192), (12, 13, 18)])
def test_live_modes_by_enumeration(grid):
    """live() against a count of the (kx, kz) modes kept by 2/3 dealiasing
    (|kx| <= Nx/3 - 1, 0 <= kz <= Nz/3 - 1, flowfield.h:578-584) times Ny;
    the tau solve's algorithmic bytes are 22 such component arrays (u_n..u_{n-2}
    and f_n..f_{n-2} read, u and q written)."""
    Nx, Ny, Nz = grid
    kept = sum(1 for kx in range(-(Nx // 2) + 1, Nx // 2) for kz in range(Nz // 2)
               if abs(kx) <= Nx // 3 - 1 and kz <= Nz // 3 - 1)
    assert bench.live(grid) == Ny * kept
    assert bench.kernel_bytes(grid, "SkewSymmetric")["k_tau_solve"] == 22 * 16 * Ny * kept


def test_serve_request_stream():
    a = bench.serve_requests(96, 4242)
    assert a == bench.serve_requests(96, 4242)
    grids = {(32, 33, 32), (48, 49, 48), (64, 65, 64), (96, 97, 96), (128, 129, 128)}
    assert {q["grid"] for q in a} == grids
    for q in a:
        Nx = q["grid"][0]
        fl = q["flags"]
        assert 100 <= q["horizon"] <= 2000
        assert 1e-3 <= q["norm"] <= 0.3
        if fl["nonlinearity"] == "SkewSymmetric":  # Couette, C1 flags
            assert 300 <= 1 / fl["nu"] <= 500
            assert fl["dt"] == pytest.approx(0.02 * 32 / Nx)
        else:  # Poiseuille, rotational, bulk velocity 1
            assert fl["nonlinearity"] == "Rotational"
            assert 1500 <= 1 / fl["nu"] <= 3000
            assert fl["dPdx"] == pytest.approx(-3 * fl["nu"])
            assert fl["dt"] == pytest.approx(0.01 * 32 / Nx)


def _counters(path, counter, rows):
    with open(path, "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n")
        for i, (name, kib) in enumerate(rows):
            f.write(f'{i},"{name}",{counter},{kib}\n')


@pytest.mark.parametrize("fetch_ratio", [0.5, 0.4999, 1.0])
def test_traffic_from_counters(tmp_path, fetch_ratio):
    """Counters in KiB per dispatch, scaled by what the same counter reads on
    the calibration copies (FETCH_SIZE has read about half a copy's bytes);
    small state-upload copies do not count as calibration, and kernel
    variants count as their step slot."""
    cal = bench.COPY_CAL_BYTES / 1024.0
    tau3 = "void cfg::k_tau_solve_wave<3, 16, 16>(cfg::Geom, double const*, cfg::TauMember const*)"
    tau2 = "void cfg::k_tau_solve_wave<2, 16, 16>(cfg::Geom, double const*, cfg::TauMember const*)"
    zw = "void cfg::k_zwave<3, 64>(cfg::Geom, HIP_vector_type<double, 2u> const*)"
    small = ("__amd_rocclr_copyBuffer", 3.0)
    fetch = [("__amd_rocclr_copyBuffer", fetch_ratio * cal)] * 3 + [small, (tau3, 100.0), (tau2, 300.0), (zw, 50.0)]
    write = [("__amd_rocclr_copyBuffer", cal)] * 3 + [small, (tau3, 10.0), (tau2, 30.0), (zw, 20.0)]
    f, w, out = tmp_path / "f.csv", tmp_path / "w.csv", tmp_path / "t.json"
    _counters(f, "FETCH_SIZE", fetch)
    _counters(w, "WRITE_SIZE", write)
    bench.traffic_from_counters(str(f), str(w), str(out), "test")
    d = json.load(open(out))
    sf = 1.0 / fetch_ratio
    assert d["_scale"]["fetch"] == pytest.approx(sf)
    assert d["_scale"]["write"] == pytest.approx(1.0)
    assert d["k_tau_solve"] == pytest.approx(1024 * (sf * 200.0 + 20.0))
    assert d["k_zpass"] == pytest.approx(1024 * (sf * 50.0 + 20.0))


class _FakeBatch:
    """Graph and tick counters of a batch whose first `eager` ticks run eagerly,
    the next `captures` ticks capture a whole-batch graph, and every later tick
    replays one."""

    def __init__(self, eager, captures):
        self.eager, self.cap, self.ticks, self.c, self.r = eager, captures, 0, 0, 0

    def advance(self, n):
        for _ in range(n):
            self.ticks += 1
            if self.ticks > self.eager + self.cap:
                self.r += 1
            elif self.ticks > self.eager:
                self.c += 1

    def graph_stats(self):
        return self.c, self.r

    def tick_stats(self):
        return self.ticks, self.c + self.r, 0.0


@pytest.mark.parametrize("eager,captures,expect", [(0, 0, 6), (4, 3, 18), (10, 9, 30)])
def test_steady_state_ends_on_a_replay_only_block(eager, captures, expect):
    """The bench's warm-up extension stops after the first block of steps that
    captured nothing and replayed one graph per tick."""
    b = _FakeBatch(eager, captures)
    assert bench.steady_state(b, 1) == expect


def test_steady_state_gives_up():
    with pytest.raises(SystemExit):
        bench.steady_state(_FakeBatch(1000, 0), 1)


class _FakeCfa:
    """What bench.FiniteAdvance uses of channelflow_amd, for a fake batch."""
    ERR_NONFINITE = -3

    class CfgError(Exception):
        def __init__(self, msg, code):
            super().__init__(msg)
            self.code = code

    @staticmethod
    def DNSFlags(**kw):
        return kw

    @staticmethod
    def FlowField(*a, **kw):
        return None


class _DivergingBatch:
    """Members 0..n-1; member `bad` stops being finite once `at` steps have run."""

    def __init__(self, n, bad, at):
        self.alive, self.bad, self.at, self.steps = set(), bad, at, 0

    def add(self, fields, flags):
        k = len(self.alive) + 100
        self.alive.add(k)
        return k

    def remove(self, k):
        self.alive.remove(k)

    def finite(self, k):
        return not (k == self.bad + 100 and self.steps >= self.at)

    def advance(self, n):
        self.steps += n
        if any(not self.finite(k) for k in self.alive):
            raise _FakeCfa.CfgError("velocity not finite after advance", _FakeCfa.ERR_NONFINITE)


def test_finite_advance_stops_diverging_member(monkeypatch):
    """A member found not finite after an advance is removed and recorded with
    the step count after which it was found; the others keep stepping and are
    the ones counted."""
    monkeypatch.setattr(bench, "member_field", lambda m: None)
    members = [dict(grid=(32, 33, 32), flags={}, norm=0.1 * (i + 1)) for i in range(4)]
    b = _DivergingBatch(4, bad=2, at=30)
    adv = bench.FiniteAdvance(_FakeCfa, b, members)
    for i in range(4):
        adv.add(i)
    assert adv(20) == [] and sorted(adv.ids) == [0, 1, 2, 3]
    assert adv(20) == [2]
    assert sorted(adv.ids) == [0, 1, 3] and adv.steps == 40
    assert adv.dropped == [dict(member=2, grid="32x33x32", norm=0.3, nonfinite_after_step=40)]
    assert adv(20) == [] and adv.points() == 3 * 32 * 33 * 32


def test_finite_advance_passes_other_errors(monkeypatch):
    monkeypatch.setattr(bench, "member_field", lambda m: None)

    class _Failing(_DivergingBatch):
        def advance(self, n):
            raise _FakeCfa.CfgError("device error", -2)

    adv = bench.FiniteAdvance(_FakeCfa, _Failing(1, 0, 0), [dict(grid=(16, 17, 16), flags={}, norm=0.1)])
    adv.add(0)
    with pytest.raises(_FakeCfa.CfgError):
        adv(5)
