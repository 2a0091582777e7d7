"""Edge-tracking batch client: host-side parts on the CPU.

Ecf (diffops.cpp:3885-3931) and L2Norm (diffops.cpp:415-487) are pinned by
closed forms on single Chebyshev modes (the y-integral of T_0^2 over [-1, 1]
is 2, of T_1^2 is 2/3; kz > 0 modes count twice) and against the oracle's
L2Norm2; the classification (edgetracking.cpp:836-893) by its branches; the
bisection's choice of the new uL, uH (edgetracking.cpp:376-446) with the
integrations replaced by a stand-in classifier.
"""
import numpy as np
import pytest

import cfref as R
from channelflow_amd import edge as E
from channelflow_amd.fields import random_field

NX, NY, NZ = 12, 9, 12


def field_v00(c, mode=0, kz=0):
    u = np.zeros((3, NY, NX, NZ // 2 + 1), dtype=np.complex128)
    u[1, mode, 0, kz] = c
    return u


def test_ecf_l2norm_closed_forms():
    c = 0.3
    assert E.ecf(field_v00(c)) == pytest.approx(2 * c * c, rel=1e-15)          # int_{-1}^{1} c^2
    assert E.ecf(field_v00(c, mode=1)) == pytest.approx(2 * c * c / 3, rel=1e-15)  # int y^2 c^2
    assert E.ecf(field_v00(c, kz=1)) == pytest.approx(4 * c * c, rel=1e-15)    # kz > 0: both signs
    assert E.ecf(field_v00(c), a=0.0, b=4.0) == pytest.approx(4 * c * c, rel=1e-15)  # (b - a) c^2
    assert E.l2norm(field_v00(c)) == pytest.approx(c, rel=1e-15)               # normalised: sqrt(mean)
    u = field_v00(c)
    u[0, 0, 0, 0] = 5.0  # u does not enter Ecf
    assert E.ecf(u) == pytest.approx(2 * c * c, rel=1e-15)


def test_ecf_matches_oracle_norm():
    g = R.Grid(NX, NY, NZ, 1.0, 1.0)
    u = random_field(NX, NY, NZ, 1.0, 1.0, norm=0.2, decay=0.5, seed=5)
    assert E.ecf(u) == pytest.approx(2 * R.l2norm2(g, u[1:3]), rel=1e-13)
    assert E.l2norm(u) == pytest.approx(np.sqrt(R.l2norm2(g, u)), rel=1e-13)


def test_classify_branches():
    af = E.AttractorFlags(EcfH=1e-2, EcfL=1e-4)
    big, small, mid = field_v00(0.2), field_v00(0.001), field_v00(0.02)   # Ecf 8e-2, 2e-6, 8e-4
    assert E.classify(big, 0, 0.0, af)[0] == E.TURBULENT
    assert E.classify(small, 0, 0.0, af)[0] == E.LAMINAR
    assert E.classify(mid, 0, 0.0, af)[0] == E.NO_ATTRACTOR
    # L2 thresholds alone, and normToHeight (Ecf / (b - a), L2 / sqrt(b - a))
    afl = E.AttractorFlags(L2H=0.1, L2L=0.01)
    assert E.classify(big, 0, 0.0, afl)[0] == E.TURBULENT
    assert E.classify(small, 0, 0.0, afl)[0] == E.LAMINAR
    afn = E.AttractorFlags(EcfH=5e-2, EcfL=1e-4, normToHeight=True)
    assert E.classify(big, 0, 0.0, afn)[0] == E.NO_ATTRACTOR   # 8e-2 / 2 < 5e-2
    # monotone decay for more than tMaxDecay time units counts as laminar
    afd = E.AttractorFlags(EcfH=1e-2, EcfL=1e-5, tMaxDecay=2)
    a, t = E.classify(mid, 0, 1.0, afd)
    assert (a, t) == (E.NO_ATTRACTOR, 1)
    a, t = E.classify(mid, 2, 1.0, afd)
    assert (a, t) == (E.LAMINAR, 3)
    a, t = E.classify(mid, 2, 0.0, afd)  # grew: the count restarts
    assert (a, t) == (E.NO_ATTRACTOR, 0)
    with pytest.raises(ValueError):
        E.classify(mid, 0, 0.0, E.AttractorFlags())
    with pytest.raises(ValueError):  # thresholds that make a field both
        E.classify(mid, 0, 0.0, E.AttractorFlags(EcfH=1e-5, EcfL=1e-2))


class StubTracker(E.EdgeTracker):
    """Attractor of a blend decided by its amplitude, no integration."""

    def __init__(self, cut):
        self.cut, self.calls = cut, []

    def attractors(self, us):
        self.calls.append(len(us))
        return [((E.TURBULENT if E.l2norm(u) > self.cut else E.LAMINAR), 1, u) for u in us]


def test_bisection_round_choices():
    uL, uH = field_v00(0.0 + 0.01), field_v00(1.0)
    # nTraj = 1: midpoint (uH + uL)/2
    tr = StubTracker(cut=0.3)
    L, H, d = tr.bisection_round(uL, uH, 1)
    assert np.array_equal(H, (uH + uL) * 0.5) and np.array_equal(L, uL)
    # nTraj = 4: blends alpha = j/5; the stub makes j = 2 (alpha 0.4) the first turbulent
    tr = StubTracker(cut=0.3)
    L, H, d = tr.bisection_round(uL, uH, 4)

    def blend(j):
        al = j * (1.0 / 5)  # alpha = j * alphaStep, as the reference computes it
        t = uL.copy()
        t *= (1 - al) / al
        t += uH
        t *= al
        return t

    assert tr.calls == [4]
    assert np.array_equal(H, blend(2)) and np.array_equal(L, uL)  # j* = 2 < 3: uL stays
    tr = StubTracker(cut=0.65)  # first turbulent j = 4: uL = blend 2
    L, H, d = tr.bisection_round(uL, uH, 4)
    assert np.array_equal(H, blend(4)) and np.array_equal(L, blend(2))
    assert d == pytest.approx(E.rel_l2_diff(L, H))
    tr = StubTracker(cut=0.9)  # none in the first 4: the next 11 blends are tried
    L, H, d = tr.bisection_round(uL, uH, 4)
    assert tr.calls == [4, 11] and np.array_equal(H, blend(5)) and np.array_equal(L, blend(3))
