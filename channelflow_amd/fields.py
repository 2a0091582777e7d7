"""Seeded synthetic initial fields for the DNS workloads (host-side numpy).

A field is a complex array of shape (Nd, Ny, Nx, Nz//2+1) in channelflow's
serial spectral layout (channelflow/flowfield.h:387-401): Chebyshev
coefficients in y, Fourier in x and z, kx = mx for mx <= Nx/2 else mx - Nx.

`random_field` builds a divergence-free, no-slip perturbation exactly in the
discrete Chebyshev basis: for each live mode (kx, kz) != (0, 0)
    v   = (1 - y^2)^2 p(y)          so v = dv/dy = 0 at y = +-1,
    eta = (1 - y^2) q(y)            (wall-normal vorticity),
    u   =  i (alpha v' - gamma eta) / kappa^2,
    w   =  i (gamma v' + alpha eta) / kappa^2,
with alpha = 2 pi kx/Lx, gamma = 2 pi kz/Lz, which satisfies
i alpha u + v' + i gamma w = 0 exactly; for (0, 0), v = 0 and u, w =
(1 - y^2) r(y) with real r. The coefficients of p, q, r are complex Gaussians
scaled by decay^(|kx| + |kz| + n/4) (n the Chebyshev index), drawn from a
counter-based generator keyed by (seed, member, slot, kx, kz, n) so that every
platform produces the same array. The kz = 0 plane is made Hermitian, aliased
and Nyquist modes are zero, and the field is scaled to the requested L2Norm
(channelflow/diffops.cpp:425-487 normalisation).
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x.copy()
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def _key(seed, member, slot, kx, kz, n, stream):
    with np.errstate(over="ignore"):
        h = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(member))
        for v in (slot, kx + 4096, kz + 4096, n, stream):
            h = _splitmix64(h ^ np.asarray(v, dtype=np.uint64))
    return h


def _gauss(seed, member, slot, kx, kz, n):
    """Complex N(0,1) + i N(0,1) per (slot, kx, kz, n) via Box-Muller."""
    u1 = (_key(seed, member, slot, kx, kz, n, 1) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u2 = (_key(seed, member, slot, kx, kz, n, 2) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)


def cheb_diff(c, L=2.0):
    """d/dy on Chebyshev coefficients along the last axis (chebyshev.cpp:672-697)."""
    c = np.asarray(c)
    N = c.shape[-1]
    d = np.zeros_like(c)
    s = 4.0 / L
    d[..., N - 2] = s * (N - 1) * c[..., N - 1]
    for n in range(N - 3, -1, -1):
        d[..., n] = d[..., n + 2] + s * (n + 1) * c[..., n + 1]
    d[..., 0] *= 0.5
    return d


def _mul_y2_exact(c):
    N = c.shape[-1]
    out = np.zeros(c.shape[:-1] + (N + 2,), dtype=c.dtype)
    for n in range(N):
        cn = c[..., n]
        # y T_n = (T_{n+1} + T_{|n-1|}) / 2, T_{-k} = T_k
        tmp = {}
        for a, w1 in ((n + 1, 0.5), (abs(n - 1), 0.5)):
            for b, w2 in ((a + 1, 0.5), (abs(a - 1), 0.5)):
                tmp[b] = tmp.get(b, 0.0) + w1 * w2
        for b, w in tmp.items():
            out[..., b] += w * cn
    return out


def one_minus_y2(c):
    c = np.asarray(c)
    y2 = _mul_y2_exact(c)
    out = -y2
    out[..., :c.shape[-1]] += c
    return out


def l2_weights(Ny):
    """W with L2Norm2(ChebyCoeff c) = c^T W c (chebyshev.cpp:758-773, normalised)."""
    W = np.zeros((Ny, Ny))
    for m in range(Ny):
        for n in range(m % 2, Ny, 2):
            W[m, n] = (1.0 - m * m - n * n) / ((1 + m - n) * (1 - m + n) * (1 + m + n) * (1 - m - n))
    return W


def l2norm2(u, live=None):
    """L2Norm2 of a spectral field (diffops.cpp:425-487) over kx in [-kxd, kxd], 0 <= kz <= kzd."""
    nd, Ny, Nx, Mz = u.shape
    W = l2_weights(Ny)
    Nz = 2 * (Mz - 1)
    kxd, kzd = live if live is not None else (Nx // 3 - 1, Nz // 3 - 1)
    mx = np.arange(-kxd, kxd + 1) % Nx
    p = u[:, :, mx, :kzd + 1]                      # (nd, Ny, nkx, nkz)
    cz = np.where(np.arange(kzd + 1) == 0, 1.0, 2.0)
    e = np.einsum("imxz,mn,inxz->xz", p.real, W, p.real) + np.einsum("imxz,mn,inxz->xz", p.imag, W, p.imag)
    return float((e * cz[None, :]).sum())


def random_field(Nx, Ny, Nz, Lx, Lz, norm=0.1, decay=0.3, seed=0, member=0, dealias=True, meanflow=True):
    """Divergence-free no-slip perturbation, shape (3, Ny, Nx, Nz//2+1)."""
    Mz = Nz // 2 + 1
    u = np.zeros((3, Ny, Nx, Mz), dtype=np.complex128)
    kxd = Nx // 3 - 1 if dealias else Nx // 2 - 1
    kzd = Nz // 3 - 1 if dealias else Nz // 2 - 1
    nv = Ny - 4      # coefficients of p: (1-y^2)^2 p has degree < Ny
    ne = Ny - 2      # coefficients of q and r
    KX, KZ = np.meshgrid(np.arange(-kxd, kxd + 1), np.arange(0, kzd + 1), indexing="ij")
    keep = ~((KZ == 0) & (KX <= 0))             # kz = 0, kx < 0 follow by conjugation; (0,0) below
    kx = KX[keep][:, None]
    kz = KZ[keep][:, None]
    scale_k = decay ** (np.abs(kx) + np.abs(kz))
    n = np.arange(nv)[None, :]
    p = _gauss(seed, member, 0, kx, kz, n) * scale_k * decay ** (n / 4.0)
    n2 = np.arange(ne)[None, :]
    q = _gauss(seed, member, 1, kx, kz, n2) * scale_k * decay ** (n2 / 4.0)
    v = one_minus_y2(one_minus_y2(p))[:, :Ny]
    eta = one_minus_y2(q)[:, :Ny]
    dv = cheb_diff(v, 2.0)
    alpha = 2 * np.pi * kx / Lx
    gamma = 2 * np.pi * kz / Lz
    k2 = alpha * alpha + gamma * gamma
    uu = 1j * (alpha * dv - gamma * eta) / k2
    ww = 1j * (gamma * dv + alpha * eta) / k2
    mx = (kx[:, 0] % Nx)
    mz = kz[:, 0]
    for comp, val in ((0, uu), (1, v), (2, ww)):
        u[comp][:, mx, mz] = val.T
    z0 = mz == 0                                # Hermitian kz = 0 plane
    mxc = (-kx[z0, 0]) % Nx
    for comp, val in ((0, uu), (1, v), (2, ww)):
        u[comp][:, mxc, 0] = np.conj(val[z0]).T
    if meanflow:
        r0 = _gauss(seed, member, 3, 0, 0, np.arange(ne)).real * decay ** (np.arange(ne) / 4.0)
        r2 = _gauss(seed, member, 4, 0, 0, np.arange(ne)).real * decay ** (np.arange(ne) / 4.0)
        u[0, :, 0, 0] = one_minus_y2(r0)[:Ny]
        u[2, :, 0, 0] = one_minus_y2(r2)[:Ny]
    nrm = np.sqrt(l2norm2(u, live=(kxd, kzd)))
    if nrm > 0 and norm is not None:
        u *= norm / nrm
    return u


def divergence(u, Lx, Lz):
    """Spectral divergence i alpha u + v' + i gamma w (Chebyshev coefficients)."""
    nd, Ny, Nx, Mz = u.shape
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)
    kz = np.arange(Mz)
    alpha = (2 * np.pi * kx / Lx)[None, :, None]
    gamma = (2 * np.pi * kz / Lz)[None, None, :]
    dv = np.moveaxis(cheb_diff(np.moveaxis(u[1], 0, -1)), -1, 0)
    return 1j * alpha * u[0] + dv + 1j * gamma * u[2]
