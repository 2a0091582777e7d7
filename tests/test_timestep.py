"""Host-side time-step control (channelflow_amd.TimeStep, after
channelflow's TimeStep, dnsflags.cpp:736-970), checked by the relations the
reference documents: n dt = dT with integer n, dtmin <= dt <= dtmax, a CFL
outside [CFLmin, CFLmax] moves to mid-range through CFL' = CFL dt'/dt, and
N dT = T after adjust_for_T."""
import pytest

import channelflow_amd as cfa


def test_construct_divides_dT():
    t = cfa.TimeStep(0.03, 0.001, 0.2, 1.0, 0.4, 0.6)
    assert t.n() == 33 and t.dt() == pytest.approx(1.0 / 33)  # round(1/0.03) = 33
    assert t.n() * t.dt() == pytest.approx(t.dT())
    with pytest.raises(ValueError):
        cfa.TimeStep(0.004, 0.005, 0.2, 1.0, 0.4, 0.6)  # dt < dtmin
    with pytest.raises(ValueError):
        cfa.TimeStep(0.02, 0.001, 0.2, 1.0, 0.7, 0.6)   # CFLmax < CFLmin


def test_adjust_to_middle():
    t = cfa.TimeStep(0.02, 0.001, 0.2, 1.0, 0.4, 0.6)
    assert t.n() == 50
    assert not t.adjust(0.5)                      # inside the band: unchanged
    assert t.adjust(0.9)                          # n' = round(2 * 50 * 0.9 / 1.0) = 90
    assert t.n() == 90 and t.dt() == pytest.approx(1 / 90)
    assert t.CFL() == pytest.approx(0.9 * (1 / 90) / 0.02)   # lands mid-band, 0.5
    assert t.adjust(0.2)                          # n' = round(2 * 90 * 0.2) = 36
    assert t.n() == 36
    fixed = cfa.TimeStep(0.02, 0.001, 0.2, 1.0, 0.4, 0.6, variable=False)
    assert not fixed.adjust(0.9) and fixed.n() == 50


def test_adjust_respects_dt_bounds():
    t = cfa.TimeStep(0.02, 0.01, 0.05, 1.0, 0.4, 0.6)
    assert t.adjust(0.01)                 # wants n = 1 (dt = 1) -> capped at dtmax: n = 20
    assert t.dt() <= 0.05 + 1e-15 and t.n() == 20
    assert t.adjust(10.0)                 # wants n = 400 (dt 0.0025) -> raised to dtmin: n = 100
    assert t.dt() >= 0.01 - 1e-15 and t.n() == 100


def test_adjust_for_T_and_desired():
    t = cfa.TimeStep(0.02, 0.001, 0.2, 1.0, 0.4, 0.6)
    t.adjust_for_T(7.3)                   # N = round(7.3) = 7, dT = 7.3/7, n = round(dT/0.02) = 52
    assert t.N() == 7 and t.dT() == pytest.approx(7.3 / 7) and t.n() == 52
    assert t.N() * t.dT() == pytest.approx(7.3) and t.n() * t.dt() == pytest.approx(t.dT())
    t = cfa.TimeStep(0.02, 0.001, 0.2, 1.0, 0.4, 0.6)
    assert t.adjust(3.0, 2.0)             # a = 3 over a_max = 2: n' = round(50 * 3 / 2) = 75
    assert t.n() == 75
    assert not t.adjust(1.0, 2.0)


def test_flags_constructor():
    f = cfa.DNSFlags(dt=0.02, dT=2.0, dtmin=0.005, dtmax=0.05, CFLmin=0.3, CFLmax=0.5)
    t = cfa.TimeStep(flags=f)
    assert t.n() == 100 and t.variable() and t.CFLmin() == 0.3 and t.CFLmax() == 0.5
