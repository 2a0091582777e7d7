"""Batch client of the DNS (multiple shooting): host-side parts on the CPU.

field2vector (channelflow/flowfield.cpp:4448-4563) is pinned by the closed-form
length the reference's comment gives, 4 (Kx + 2 Kx Kz + Kz)(Ny - 3) + 2 (Ny - 2),
and by the positions of tagged coefficients; the step size eps by
DSI::Jacobian's rule (nsolver/dsi.cpp:68-69).
"""
import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.dsi import MultiShooting, fd_step, field2vector, field2vector_norm2, field2vector_size
from channelflow_amd.timestep import TimeStep


@pytest.mark.parametrize("grid", [(32, 33, 32), (32, 49, 32), (48, 49, 48), (24, 17, 12), (16, 9, 16)])
def test_field2vector_size_closed_form(grid):
    Nx, Ny, Nz = grid
    Kx, Kz = Nx // 3 - 1, Nz // 3 - 1
    assert field2vector_size(Nx, Ny, Nz) == 4 * (Kx + 2 * Kx * Kz + Kz) * (Ny - 3) + 2 * (Ny - 2)
    u = np.zeros((3, Ny, Nx, Nz // 2 + 1), dtype=np.complex128)
    assert field2vector(u).size == field2vector_size(Nx, Ny, Nz)


def test_field2vector_order():
    """Each coefficient tagged with its own (i, ny, mx, mz); the vector must
    list them as the reference's loops do."""
    Nx, Ny, Nz = 12, 9, 12
    Kx, Kz = Nx // 3 - 1, Nz // 3 - 1
    u = np.zeros((3, Ny, Nx, Nz // 2 + 1), dtype=np.complex128)
    tag = lambda i, ny, mx, mz: 1000 * i + 100 * ny + 10 * mx + mz  # noqa: E731
    for i in range(3):
        for ny in range(Ny):
            for mx in range(Nx):
                for mz in range(Nz // 2 + 1):
                    u[i, ny, mx, mz] = tag(i, ny, mx, mz) + 0.5j * tag(i, ny, mx, mz)
    expect = []
    expect += [tag(0, ny, 0, 0) for ny in range(2, Ny)]
    expect += [tag(2, ny, 0, 0) for ny in range(2, Ny)]
    for kx in range(1, Kx + 1):
        for ny in range(2, Ny):
            expect += [tag(2, ny, kx, 0), 0.5 * tag(2, ny, kx, 0)]
        for ny in range(3, Ny - 1):
            expect += [tag(0, ny, kx, 0), 0.5 * tag(0, ny, kx, 0)]
    for kz in range(1, Kz + 1):
        for ny in range(2, Ny):
            expect += [tag(0, ny, 0, kz), 0.5 * tag(0, ny, 0, kz)]
        for ny in range(3, Ny - 1):
            expect += [tag(2, ny, 0, kz), 0.5 * tag(2, ny, 0, kz)]
    for kx in list(range(-Kx, 0)) + list(range(1, Kx + 1)):
        mx = kx % Nx
        for kz in range(1, Kz + 1):
            for ny in range(2, Ny):
                expect += [tag(0, ny, mx, kz), 0.5 * tag(0, ny, mx, kz)]
            for ny in range(3, Ny - 1):
                expect += [tag(2, ny, mx, kz), 0.5 * tag(2, ny, mx, kz)]
    assert np.array_equal(field2vector(u), np.array(expect, dtype=np.float64))


@pytest.mark.parametrize("grid", [(12, 9, 12), (32, 49, 32), (6, 7, 6)])
def test_field2vector_norm2(grid):
    Nx, Ny, Nz = grid
    rng = np.random.default_rng(7)
    u = rng.standard_normal((3, Ny, Nx, Nz // 2 + 1)) + 1j * rng.standard_normal((3, Ny, Nx, Nz // 2 + 1))
    v = field2vector(u)
    assert field2vector_norm2(u) == pytest.approx(float(np.dot(v, v)), rel=1e-13)


def test_fd_step_rule():
    Nx, Ny, Nz = 12, 9, 12
    rng = np.random.default_rng(3)
    dx = [rng.standard_normal((3, Ny, Nx, Nz // 2 + 1)) + 1j * rng.standard_normal((3, Ny, Nx, Nz // 2 + 1))
          for _ in range(2)]
    mag = np.sqrt(sum(np.sum(field2vector(d) ** 2) for d in dx))
    assert fd_step(dx, 1e-7) == pytest.approx(1e-7 / mag, rel=1e-13)
    small = [1e-12 * d for d in dx]
    assert fd_step(small, 1e-7) == 1.0  # |dx| < epsDx: eps = 1


def test_multishooting_time_fit():
    """f^T's step count and dt are TimeStep::adjust_for_T's (cfdsi.cpp:705-706)."""
    for dt, T in [(0.02, 4.0), (0.02, 0.3), (0.03125, 1.0), (0.01, 0.07)]:
        fl = cfa.DNSFlags(dt=dt, variabledt=False)
        ms = MultiShooting(16, 17, 16, 1.0, 1.0, fl, T)
        ts = TimeStep(flags=fl)
        ts.adjust_for_T(T)
        assert ms.nsteps == ts.N() * ts.n() and ms.flags.dt == ts.dt()
        assert abs(ms.nsteps * ms.flags.dt - T) <= 1e-12 * T
    # variable dt: the same fit, and the integrations then adapt dt per dT (cfdsi.cpp:733-767)
    ms = MultiShooting(16, 17, 16, 1.0, 1.0, cfa.DNSFlags(dt=0.02, variabledt=True), 1.0)
    assert ms.variable and ms.timestep.N() * ms.timestep.dT() == pytest.approx(1.0, rel=1e-14)
    with pytest.raises(ValueError):
        MultiShooting(16, 17, 16, 1.0, 1.0, cfa.DNSFlags(dt=0.02), -1.0)
