// Per-mode FP64 math of the DNS time step, written once for the device
// kernels and the host-side unit checks.
//
// Every per-mode vector is a strided view into a "mode-fastest" array:
// element n of mode m lives at base[n * stride + m] (real arrays) or, for the
// interleaved complex state, at base[2 * (n * M + m) + part]. Threads that own
// consecutive modes therefore touch consecutive addresses at every step of the
// Chebyshev recurrences below.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CFG_HD __host__ __device__ inline
#else
#define CFG_HD inline
#endif

// The recurrences below are evaluated without fused multiply-add contraction,
// so each operation rounds as in the reference's x86-64 build: the tau-
// corrected pressure amplifies per-operation rounding differences. The pragma
// is lexical and ends with this header (contraction is on again below it), so
// the kernels that include it keep their fused multiply-adds.
#pragma clang fp contract(off)

namespace cfg {

struct SV {  // strided view of one Chebyshev profile
    double* p;
    long s;
    CFG_HD double& operator[](int n) const { return p[(long)n * s]; }
};

// ---------------------------------------------------------------------------
// Chebyshev primitives on [a,b] (channelflow/chebyshev.cpp:405-512, 672-697).
// ---------------------------------------------------------------------------

// du = d/dy u by the backward three-term recurrence, du[0] halved. The two
// newest terms are carried in registers, so no step waits on a store.
CFG_HD void cheb_diff(SV u, SV du, int N, double L) {
    const int Nb = N - 1;
    const double scale = 4.0 / L;
    double d2 = 0.0, d1 = scale * Nb * u[Nb];
    du[Nb] = d2;
    du[Nb - 1] = d1;
#pragma unroll 4
    for (int n = Nb - 2; n > 0; --n) {
        const double dn = d2 + scale * (n + 1) * u[n + 1];
        du[n] = dn;
        d2 = d1;
        d1 = dn;
    }
    du[0] = (d2 + scale * 1 * u[1]) * 0.5;
}

// Values of d/dy u at y=b and y=a, summed from the highest mode down exactly
// as eval_b()/eval_a() sum the stored derivative. The derivative itself is not
// stored; du is unused scratch kept for the call signature.
CFG_HD void cheb_diff_ends(SV u, SV du, int N, double L, double& at_b, double& at_a) {
    const int Nb = N - 1;
    const double scale = 4.0 / L;
    double d2 = 0.0, d1 = scale * Nb * u[Nb];
    double sb = 0.0, sa = 0.0;
    sb += d2;
    sa += d2 * ((Nb % 2 == 0) ? 1 : -1);
    sb += d1;
    sa += d1 * (((Nb - 1) % 2 == 0) ? 1 : -1);
#pragma unroll 4
    for (int n = Nb - 2; n > 0; --n) {
        const double dn = d2 + scale * (n + 1) * u[n + 1];
        sb += dn;
        sa += dn * ((n % 2 == 0) ? 1 : -1);
        d2 = d1;
        d1 = dn;
    }
    const double d0 = (d2 + scale * 1 * u[1]) * 0.5;
    sb += d0;
    sa += d0;
    at_b = sb;
    at_a = sa;
}

CFG_HD double cheb_eval_b(SV u, int N) {
    double s = 0.0;
    for (int n = N - 1; n >= 0; --n) s += u[n];
    return s;
}

CFG_HD double cheb_eval_a(SV u, int N) {
    double s = 0.0;
    for (int n = N - 1; n >= 0; --n) s += u[n] * ((n % 2 == 0) ? 1 : -1);
    return s;
}

CFG_HD double cheb_mean(SV u, int N) {
    double s = u[0];
    for (int n = 2; n < N; n += 2) s -= u[n] / (double)(n * n - 1);
    return s;
}

// ---------------------------------------------------------------------------
// Chebyshev-tau Helmholtz operator  nu u'' - lambda u = f,  u(a)=ua, u(b)=ub
// (Canuto & Hussaini 5.1.24, channelflow/helmholtz.cpp:18-95). Even and odd
// Chebyshev modes decouple into two "banded tridiagonal" systems: a full row 0
// plus a tridiagonal, factored UL without pivoting
// (channelflow/bandedtridiag.cpp:212-277).
//
// Factor storage is packed by Chebyshev index n (even n -> even system row
// n/2, odd n -> odd system row (n-1)/2):
//   band[n] : row-0 entry of column n's row after factoring; band[0], band[1]
//             are diag(0) of the even and odd systems,
//   up[n]   : factored super-diagonal of row n (n >= 2),
//   invd[n] : 1/diag of row n after factoring (n >= 2).
// The sub-diagonal is not stored: factoring leaves it unchanged, and it is
// recomputed from lambda below with the same arithmetic.
// ---------------------------------------------------------------------------
struct HelmholtzView {
    SV band, up, invd;
    double lambda;
    int N;  // number of Chebyshev modes (odd)
};

CFG_HD int helm_c(int n, int Nc) { return (n == 0 || n == Nc) ? 2 : 1; }   // helmholtz.h:66
CFG_HD int helm_beta(int n, int Nc) { return (n > Nc - 2) ? 0 : 1; }       // helmholtz.h:65

// sub-diagonal of the A matrix, row with Chebyshev index n >= 2
CFG_HD double helm_lo(int n, double lambda, int Nc) { return -(helm_c(n - 2, Nc) * lambda) / (4 * n * (n - 1)); }

// Build and UL-factor both parity systems for (lambda, nu) into h.
CFG_HD void helm_factor(HelmholtzView h, double nu, double a, double b) {
    const int N = h.N, Nc = N - 1;
    const double lambda = h.lambda;
    const double nuscaled = nu / (((b - a) / 2) * ((b - a) / 2));
    for (int par = 0; par < 2; ++par) {
        const int M = (par == 0) ? Nc / 2 + 1 : Nc / 2;
        // assemble: band row = 1, rows i>=1 from the formulas; invd holds diag
        for (int i = 0; i < M; ++i) {
            const int n = 2 * i + par;
            h.band[n] = 1.0;
            if (i >= 1) {
                const int nn = n * n;
                h.invd[n] = (nuscaled + (helm_beta(n, Nc) * lambda) / (2 * (nn - 1)));
                h.up[n] = helm_beta(n + 2, Nc) ? -lambda / (4 * n * (n + 1)) : 0.0;
            }
        }
        // UL decomposition (bandedtridiag.cpp:212-229)
        for (int k = M - 1; k > 1; --k) {
            const int nk = 2 * k + par, nk1 = 2 * (k - 1) + par;
            const double Akk = h.invd[nk];
            const double w = helm_lo(nk, lambda, Nc);
            h.up[nk1] /= Akk;
            h.invd[nk1] -= w * h.up[nk1];
            h.band[nk] /= Akk;
            h.band[nk1] -= w * h.band[nk];
        }
        const int n1 = 2 + par;
        h.band[n1] /= h.invd[n1];
        h.band[par] -= helm_lo(n1, lambda, Nc) * h.band[n1];
        for (int i = 1; i < M; ++i) h.invd[2 * i + par] = 1.0 / h.invd[2 * i + par];
    }
}

// u = A^{-1} (B f) with u(a)=ua, u(b)=ub. u and f must not alias.
CFG_HD void helm_solve(const HelmholtzView& h, SV u, SV f, double ua, double ub) {
    const int N = h.N, Nc = N - 1;
    const double lambda = h.lambda;
    for (int par = 0; par < 2; ++par) {
        const int M = (par == 0) ? Nc / 2 + 1 : Nc / 2;
        // u = B f  (B is lambda-independent; row 0 is overwritten by the BC)
        for (int i = 1; i < M - 1; ++i) {
            const int n = 2 * i + par, nn = n * n;
            const double lo = ((double)helm_c(n - 2, Nc)) / (4 * n * (n - 1));
            const double dg = -((double)helm_beta(n, Nc)) / (2 * (nn - 1));
            const double upv = helm_beta(n + 2, Nc) ? 1.0 / (4 * n * (n + 1)) : 0.0;
            u[n] = lo * f[n - 2] + dg * f[n] + upv * f[n + 2];
        }
        {
            const int i = M - 1, n = 2 * i + par, nn = n * n;
            const double lo = ((double)helm_c(n - 2, Nc)) / (4 * n * (n - 1));
            const double dg = -((double)helm_beta(n, Nc)) / (2 * (nn - 1));
            u[n] = lo * f[n - 2] + dg * f[n];
        }
    }
    u[0] = (ub + ua) / 2;
    u[1] = (ub - ua) / 2;
    for (int par = 0; par < 2; ++par) {
        const int M = (par == 0) ? Nc / 2 + 1 : Nc / 2;
        // U y = b, backward (bandedtridiag.cpp:238-243), carrying b[i+1]
        double nxt = u[2 * (M - 1) + par];
#pragma unroll 4
        for (int i = M - 2; i > 0; --i) {
            const double v = u[2 * i + par] - h.up[2 * i + par] * nxt;
            u[2 * i + par] = v;
            nxt = v;
        }
        double acc = u[par];
#pragma unroll 4
        for (int j = 1; j < M; ++j) acc -= h.band[2 * j + par] * u[2 * j + par];
        double prev = acc / h.band[par];
        u[par] = prev;
        // L x = y, forward, carrying x[i-1]
#pragma unroll 4
        for (int i = 1; i < M; ++i) {
            const double v = (u[2 * i + par] - helm_lo(2 * i + par, lambda, Nc) * prev) * h.invd[2 * i + par];
            u[2 * i + par] = v;
            prev = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-mode tau solver (Canuto & Hussaini 7.3, channelflow/tausolver.cpp).
// ---------------------------------------------------------------------------
struct TauView {
    HelmholtzView ph;  // pressure: lambda = kappa^2, nu = 1
    HelmholtzView vh;  // velocity: lambda = lambda_t + nu kappa^2
    SV Pp, Pm, vp, vm, P0, v0;
    double* scal;  // i00 i01 i10 i11 sigma0_Nb1 sigma0_Nb, at scal[k * sstride]
    long sstride;
    double nu, a, b, kx2pi, kz2pi;  // 2 pi kx/Lx, 2 pi kz/Lz
    bool zero_mode, taucorrection;
};

CFG_HD int tau_nfunc(int k, int Nb) {  // tausolver.cpp:15
    if (k == 0) return Nb - 1;
    if (k == Nb) return 0;
    if (k % 2 == 0) return 2 * (Nb - 1);
    return 2 * Nb;
}

CFG_HD void tau_influence(const TauView& t, SV P, SV v, SV tmp) {  // tausolver.cpp:178-191
    const int N = t.vh.N;
    double dp, dm;
    cheb_diff_ends(v, tmp, N, t.b - t.a, dp, dm);
    const double* sc = t.scal;
    const long ss = t.sstride;
    const double delta_plus = -sc[0 * ss] * dp - sc[1 * ss] * dm;
    const double delta_minus = -sc[2 * ss] * dp - sc[3 * ss] * dm;
    for (int i = 0; i < N; ++i) {
        P[i] += delta_plus * t.Pp[i] + delta_minus * t.Pm[i];
        v[i] += delta_plus * t.vp[i] + delta_minus * t.vm[i];
    }
}

// TauSolver constructor (tausolver.cpp:81-176): factors both Helmholtz
// operators and the homogeneous / tau-correction solutions. s0, s1 scratch.
CFG_HD void tau_setup(const TauView& t, SV s0, SV s1) {
    const int N = t.vh.N, Nb = N - 1;
    const double L = t.b - t.a;
    helm_factor(t.ph, 1.0, t.a, t.b);
    helm_factor(t.vh, t.nu, t.a, t.b);
    double* sc = t.scal;
    const long ss = t.sstride;
    for (int n = 0; n < N; ++n) s0[n] = 0.0;
    helm_solve(t.ph, t.Pp, s0, 0.0, 1.0);
    cheb_diff(t.Pp, s1, N, L);
    helm_solve(t.vh, t.vp, s1, 0.0, 0.0);
    helm_solve(t.ph, t.Pm, s0, 1.0, 0.0);
    cheb_diff(t.Pm, s1, N, L);
    helm_solve(t.vh, t.vm, s1, 0.0, 0.0);
    if (t.zero_mode) {  // influence matrix is singular for kx=kz=0 and unused
        for (int k = 0; k < 6; ++k) sc[k * ss] = 0.0;
        for (int n = 0; n < N; ++n) { t.P0[n] = 0.0; t.v0[n] = 0.0; }
        return;
    }
    double A, B, C, D;
    cheb_diff_ends(t.vp, s1, N, L, A, C);
    cheb_diff_ends(t.vm, s1, N, L, B, D);
    const double disc = A * D - B * C;
    sc[0 * ss] = D / disc;
    sc[1 * ss] = -B / disc;
    sc[2 * ss] = -C / disc;
    sc[3 * ss] = A / disc;
    const double c = 2 / (t.b - t.a);
    for (int i = 0; i <= Nb; ++i) s0[i] = c * tau_nfunc(i, Nb);
    helm_solve(t.ph, t.P0, s0, 0.0, 0.0);
    cheb_diff(t.P0, s1, N, L);  // dP0/dy before the influence correction
    const double dP0_Nb1 = s1[Nb - 1], dP0_Nb = s1[Nb];
    helm_solve(t.vh, t.v0, s1, 0.0, 0.0);
    tau_influence(t, t.P0, t.v0, s0);
    // (v0)'' vanishes in its top two modes, so its nu-term drops out here
    sc[4 * ss] = t.vh.lambda * t.v0[Nb - 1] + dP0_Nb1 - t.nu * 0.0;
    sc[5 * ss] = t.vh.lambda * t.v0[Nb] + dP0_Nb - t.nu * 0.0;
}

// solve_P_and_v (tausolver.cpp:193-251) for one real part. r is div(R) for
// this part, Ry the y-momentum right-hand side; tmp scratch.
CFG_HD void tau_P_and_v(const TauView& t, SV P, SV v, SV r, SV Ry, SV tmp) {
    const int N = t.vh.N, Nb = N - 1;
    const double L = t.b - t.a;
    helm_solve(t.ph, P, r, 0.0, 0.0);
    if (t.zero_mode) {
        for (int i = 0; i < N; ++i) v[i] = 0.0;
        return;
    }
    cheb_diff(P, tmp, N, L);
    for (int i = 0; i < N; ++i) tmp[i] -= Ry[i];
    helm_solve(t.vh, v, tmp, 0.0, 0.0);
    tau_influence(t, P, v, tmp);
    if (!t.taucorrection) return;
    const double lambda = t.vh.lambda;
    // v'' has zero top two modes; dP/dy has dP[Nb] = 0, dP[Nb-1] = (4/L) Nb P[Nb]
    double sigma1_Nb = lambda * v[Nb] - t.nu * 0.0 - Ry[Nb];
    double sigma1_Nb1 = lambda * v[Nb - 1] - t.nu * 0.0 - Ry[Nb - 1];
    sigma1_Nb += 0.0;
    sigma1_Nb1 += (4.0 / L) * Nb * P[Nb];
    const double* sc = t.scal;
    const long ss = t.sstride;
    const double sigmaNb = sigma1_Nb / (1.0 - sc[5 * ss]);
    const double sigmaNb1 = sigma1_Nb1 / (1.0 - sc[4 * ss]);
    for (int i = 0; i <= Nb; ++i) {
        P[i] += ((i % 2 == 0) ? sigmaNb1 : sigmaNb) * t.P0[i];
        v[i] += ((i % 2 == 0) ? sigmaNb : sigmaNb1) * t.v0[i];
    }
}

// Complex right-hand sides and solutions as separate re/im strided views.
struct CV {
    SV re, im;
};

// TauSolver::solve (tausolver.cpp:347-402). s0, s1 scratch.
CFG_HD void tau_solve(const TauView& t, CV u, CV v, CV w, CV P, CV Rx, CV Ry, CV Rz, SV s0, SV s1) {
    const int N = t.vh.N;
    const double L = t.b - t.a;
    cheb_diff(Ry.re, s0, N, L);
    for (int n = 0; n < N; ++n) s0[n] -= t.kx2pi * Rx.im[n] + t.kz2pi * Rz.im[n];
    tau_P_and_v(t, P.re, v.re, s0, Ry.re, s1);
    cheb_diff(Ry.im, s0, N, L);
    for (int n = 0; n < N; ++n) s0[n] += t.kx2pi * Rx.re[n] + t.kz2pi * Rz.re[n];
    tau_P_and_v(t, P.im, v.im, s0, Ry.im, s1);
    for (int n = 0; n < N; ++n) {
        s0[n] = -(t.kx2pi * P.im[n]) - Rx.re[n];
        s1[n] = t.kx2pi * P.re[n] - Rx.im[n];
    }
    helm_solve(t.vh, u.re, s0, 0.0, 0.0);
    helm_solve(t.vh, u.im, s1, 0.0, 0.0);
    for (int n = 0; n < N; ++n) {
        s0[n] = -(t.kz2pi * P.im[n]) - Rz.re[n];
        s1[n] = t.kz2pi * P.re[n] - Rz.im[n];
    }
    helm_solve(t.vh, w.re, s0, 0.0, 0.0);
    helm_solve(t.vh, w.im, s1, 0.0, 0.0);
}

// Helmholtz with unknown constant forcing mu and mean constraint
// (helmholtz.cpp:158-213): nu u'' - lambda u = f + mu, mean(u) = umean.
CFG_HD double helm_solve_mean(const HelmholtzView& h, double nu, SV u, SV f, double umean, SV s0, SV s1) {
    const int N = h.N;
    for (int n = 0; n < N; ++n) s1[n] = f[n];
    helm_solve(h, s0, s1, 0.0, 0.0);
    const double uamean = cheb_mean(s0, N);
    for (int n = 0; n < N; ++n) s1[n] = 0.0;
    s1[0] = nu;
    helm_solve(h, s0, s1, 0.0, 0.0);
    const double ucmean = cheb_mean(s0, N);
    const double mu = nu * (umean - uamean) / ucmean;
    for (int n = 0; n < N; ++n) s1[n] = f[n];
    s1[0] += mu;
    helm_solve(h, u, s1, 0.0, 0.0);
    return mu;
}

// Bulk-velocity form for kx=kz=0 (tausolver.cpp:404-450). Needs four scratch
// vectors; returns dPdx, dPdz.
CFG_HD void tau_solve_bulk(const TauView& t, CV u, CV v, CV w, CV P, CV Rx, CV Ry, CV Rz, double umean, double wmean,
                           SV s0, SV s1, SV s2, SV s3, double& dPdx, double& dPdz) {
    const int N = t.vh.N;
    const double L = t.b - t.a;
    cheb_diff(Ry.re, s0, N, L);
    for (int n = 0; n < N; ++n) s0[n] -= t.kx2pi * Rx.im[n] + t.kz2pi * Rz.im[n];
    tau_P_and_v(t, P.re, v.re, s0, Ry.re, s1);
    cheb_diff(Ry.im, s0, N, L);
    for (int n = 0; n < N; ++n) s0[n] += t.kx2pi * Rx.re[n] + t.kz2pi * Rz.re[n];
    tau_P_and_v(t, P.im, v.im, s0, Ry.im, s1);
    for (int n = 0; n < N; ++n) {
        s0[n] = -(t.kx2pi * P.im[n]) - Rx.re[n];
        s1[n] = t.kx2pi * P.re[n] - Rx.im[n];
    }
    dPdx = helm_solve_mean(t.vh, t.nu, u.re, s0, umean, s2, s3);
    helm_solve(t.vh, u.im, s1, 0.0, 0.0);
    for (int n = 0; n < N; ++n) {
        s0[n] = -(t.kz2pi * P.im[n]) - Rz.re[n];
        s1[n] = t.kz2pi * P.re[n] - Rz.im[n];
    }
    dPdz = helm_solve_mean(t.vh, t.nu, w.re, s0, wmean, s2, s3);
    helm_solve(t.vh, w.im, s1, 0.0, 0.0);
}

}  // namespace cfg

#pragma clang fp contract(fast)
