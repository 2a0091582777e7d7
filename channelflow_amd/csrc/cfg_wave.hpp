// Wavefront-parallel Chebyshev-tau algebra for one (kx, kz) mode per wave64.
//
// A Chebyshev profile of N (odd) coefficients is held split by parity, the
// way the Chebyshev-tau Helmholtz systems decouple: E[i] = c[2i] (Me = (N+1)/2
// values) and O[i] = c[2i+1] (Mo = (N-1)/2 values), blocked K per lane, so
// lane l owns E[Kl..Kl+K-1], O[Kl..Kl+K-1], i.e. the 2K contiguous
// coefficients c[2Kl .. 2Kl+2K-1]. Entries past Me / Mo are zero.
//
// The sequential recurrences of channelflow's Chebyshev algebra become scans:
//   d/dy (chebyshev.cpp:672-697)  -> suffix sums within each parity,
//   UL back/forward substitution (bandedtridiag.cpp:232-277) -> affine
//     suffix / prefix scans, the full row 0 -> a wave reduction,
//   eval_a/eval_b/mean (chebyshev.cpp:405-512) -> wave reductions.
// The results equal the sequential ones up to summation order.
#pragma once

#include <hip/hip_runtime.h>

namespace cfg {

template <int K>
struct WV {
    double e[K], o[K];
};

struct WCtx {
    int N, Me, Mo, Nb, Nc, lane;
    double L;
};

__device__ __forceinline__ double wdn(double v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ double wup(double v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ double wbc(double v, int src) { return __shfl(v, src, 64); }

// sum over the wave, result identical on every lane (tree to lane 0, broadcast)
__device__ __forceinline__ double wsum(double v, int lane) {
    for (int d = 32; d >= 1; d >>= 1) {
        const double y = wdn(v, d);
        if (lane + d < 64) v += y;
    }
    return wbc(v, 0);
}

template <int K>
__device__ __forceinline__ WV<K> wzero() {
    WV<K> v;
    for (int k = 0; k < K; ++k) v.e[k] = v.o[k] = 0.0;
    return v;
}

// contiguous coefficients c[0..N) (stride 1)
template <int K>
__device__ __forceinline__ WV<K> wload(const double* p, const WCtx& c) {
    WV<K> v;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        v.e[k] = (i < c.Me) ? p[2 * i] : 0.0;
        v.o[k] = (i < c.Mo) ? p[2 * i + 1] : 0.0;
    }
    return v;
}

// strided coefficients c[n] at p[n * s]
template <int K>
__device__ __forceinline__ WV<K> wload_s(const double* p, long s, const WCtx& c) {
    WV<K> v;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        v.e[k] = (i < c.Me) ? p[(long)(2 * i) * s] : 0.0;
        v.o[k] = (i < c.Mo) ? p[(long)(2 * i + 1) * s] : 0.0;
    }
    return v;
}

template <int K>
__device__ __forceinline__ void wstore_s(double* p, long s, const WV<K>& v, const WCtx& c) {
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        if (i < c.Me) p[(long)(2 * i) * s] = v.e[k];
        if (i < c.Mo) p[(long)(2 * i + 1) * s] = v.o[k];
    }
}

// coefficient n of v, broadcast to all lanes
// Blocked entry i of one parity array, broadcast to all lanes. Each candidate
// passes through an empty asm operand: without it the select chain is folded
// into an indexed load, which puts the array in scratch memory.
template <int K>
__device__ __forceinline__ double wpick(const double (&a)[K], int i, const WCtx& c) {
    const int owner = i / K, slot = i - owner * K;
    double x = 0.0;
    for (int k = 0; k < K; ++k) {
        double t = a[k];
        asm volatile("" : "+v"(t));
        x = (k == slot) ? t : x;
    }
    return wbc(x, owner);
}

// coefficients N-1 (even, N odd) and N-2 (odd) of v, broadcast to all lanes
template <int K>
__device__ __forceinline__ double wtop(const WV<K>& v, const WCtx& c) { return wpick<K>(v.e, c.Me - 1, c); }
template <int K>
__device__ __forceinline__ double wtop1(const WV<K>& v, const WCtx& c) { return wpick<K>(v.o, c.Mo - 1, c); }

// a*x + b*y elementwise
template <int K>
__device__ __forceinline__ WV<K> waxpby(double a, const WV<K>& x, double b, const WV<K>& y) {
    WV<K> r;
    for (int k = 0; k < K; ++k) {
        r.e[k] = a * x.e[k] + b * y.e[k];
        r.o[k] = a * x.o[k] + b * y.o[k];
    }
    return r;
}

// suffix sums of t over the blocked index: out[i] = sum_{j >= i} t[j]
// (inclusive) and excl[i] = sum_{j > i} t[j]
template <int K>
__device__ __forceinline__ void wsuffix(const double (&t)[K], double (&incl)[K], double (&excl)[K], int lane) {
    double loc[K];
    double acc = 0.0;
    for (int k = K - 1; k >= 0; --k) {
        acc = t[k] + acc;
        loc[k] = acc;
    }
    double x = loc[0];
    for (int d = 1; d < 64; d <<= 1) {
        const double y = wdn(x, d);
        if (lane + d < 64) x += y;
    }
    double after = wdn(x, 1);
    if (lane == 63) after = 0.0;
    for (int k = 0; k < K; ++k) {
        incl[k] = loc[k] + after;
        excl[k] = (k + 1 < K) ? loc[k + 1] + after : after;
    }
}

// d/dy of Chebyshev coefficients on an interval of length L:
// du[2i] = sum_{j>=i} (4/L)(2j+1) c[2j+1],  du[2i+1] = sum_{j>i} (4/L)(2j) c[2j], du[0] halved
template <int K>
__device__ __forceinline__ WV<K> wdiff(const WV<K>& u, const WCtx& c) {
    const double s = 4.0 / c.L;
    double to[K], te[K], inc[K], exc[K];
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        to[k] = (i < c.Mo) ? s * (2 * i + 1) * u.o[k] : 0.0;
        te[k] = (i < c.Me) ? s * (2 * i) * u.e[k] : 0.0;
    }
    WV<K> d;
    wsuffix<K>(to, inc, exc, c.lane);
    for (int k = 0; k < K; ++k) d.e[k] = (K * c.lane + k < c.Me) ? inc[k] : 0.0;
    wsuffix<K>(te, inc, exc, c.lane);
    for (int k = 0; k < K; ++k) d.o[k] = (K * c.lane + k < c.Mo) ? exc[k] : 0.0;
    if (c.lane == 0) d.e[0] *= 0.5;
    return d;
}

// values at y=b (sum of coefficients) and y=a (alternating sum)
template <int K>
__device__ __forceinline__ void wends(const WV<K>& v, const WCtx& c, double& at_b, double& at_a) {
    double se = 0.0, so = 0.0;
    for (int k = 0; k < K; ++k) {
        se += v.e[k];
        so += v.o[k];
    }
    se = wsum(se, c.lane);
    so = wsum(so, c.lane);
    at_b = se + so;
    at_a = se - so;
}

// Affine suffix scan: x_i = cc_i + aa_i x_{i+1}; maps past the chain end are
// the identity (cc = 0, aa = 1) and the last real map is constant (aa = 0).
// On return cc[k] holds x_i.
template <int K>
__device__ __forceinline__ void waff_suffix(double (&cc)[K], const double (&aa)[K], int lane) {
    double Fc[K], Fa[K];
    double xc = 0.0, xa = 1.0;
    for (int k = K - 1; k >= 0; --k) {
        xc = cc[k] + aa[k] * xc;
        xa = aa[k] * xa;
        Fc[k] = xc;
        Fa[k] = xa;
    }
    for (int d = 1; d < 64; d <<= 1) {
        const double yc = wdn(xc, d), ya = wdn(xa, d);
        if (lane + d < 64) {
            xc = xc + xa * yc;
            xa = xa * ya;
        }
    }
    double ec = wdn(xc, 1);
    if (lane == 63) ec = 0.0;
    for (int k = 0; k < K; ++k) cc[k] = Fc[k] + Fa[k] * ec;
}

// Affine prefix scan: y_i = cc_i + aa_i y_{i-1}; the first map is constant.
template <int K>
__device__ __forceinline__ void waff_prefix(double (&cc)[K], const double (&aa)[K], int lane) {
    double Gc[K], Ga[K];
    double xc = 0.0, xa = 1.0;
    for (int k = 0; k < K; ++k) {
        xc = cc[k] + aa[k] * xc;
        xa = aa[k] * xa;
        Gc[k] = xc;
        Ga[k] = xa;
    }
    for (int d = 1; d < 64; d <<= 1) {
        const double yc = wup(xc, d), ya = wup(xa, d);
        if (lane >= d) {
            xc = xc + xa * yc;
            xa = xa * ya;
        }
    }
    double ec = wup(xc, 1);
    if (lane == 0) ec = 0.0;
    for (int k = 0; k < K; ++k) cc[k] = Gc[k] + Ga[k] * ec;
}

__device__ __forceinline__ int wh_c(int n, int Nc) { return (n == 0 || n == Nc) ? 2 : 1; }   // helmholtz.h:66
__device__ __forceinline__ int wh_beta(int n, int Nc) { return (n > Nc - 2) ? 0 : 1; }      // helmholtz.h:65

// One parity system of nu u'' - lambda u = f with u[0] set to bc:
// x = A^{-1} (B f) using the UL factors band/up/invd of this parity
// (helmholtz.cpp:79-95 with bandedtridiag.cpp:232-277).
template <int K>
__device__ __forceinline__ void whelm_par(const double (&f)[K], const double (&band)[K], const double (&up)[K],
                                          const double (&invd)[K], double lambda, double bc, int par, int M,
                                          const WCtx& c, double (&x)[K]) {
    const int lane = c.lane, Nc = c.Nc;
    // neighbours f[i-1], f[i+1]
    double fm[K], fp[K];
    for (int k = 0; k < K; ++k) {
        fm[k] = (k > 0) ? f[k - 1] : 0.0;
        fp[k] = (k + 1 < K) ? f[k + 1] : 0.0;
    }
    {
        const double lastp = wup(f[K - 1], 1);
        const double firstn = wdn(f[0], 1);
        if (lane > 0) fm[0] = lastp;
        if (lane < 63) fp[K - 1] = firstn;
        else fp[K - 1] = 0.0;
    }
    // u = B f (rows 1..M-1), u[0] = bc
    double cc[K], aa[K];
    for (int k = 0; k < K; ++k) {
        const int i = K * lane + k;
        const int n = 2 * i + par;
        double u = 0.0;
        if (i == 0) {
            u = bc;
        } else if (i < M) {
            const int nn = n * n;
            const double lo = ((double)wh_c(n - 2, Nc)) / (4 * n * (n - 1));
            const double dg = -((double)wh_beta(n, Nc)) / (2 * (nn - 1));
            if (i < M - 1) {
                const double upv = wh_beta(n + 2, Nc) ? 1.0 / (4 * n * (n + 1)) : 0.0;
                u = lo * fm[k] + dg * f[k] + upv * fp[k];
            } else {
                u = lo * fm[k] + dg * f[k];
            }
        }
        x[k] = u;
        // backward substitution maps on rows 1..M-1
        if (i >= 1 && i < M - 1) {
            cc[k] = u;
            aa[k] = -up[k];
        } else if (i == M - 1 && i >= 1) {
            cc[k] = u;
            aa[k] = 0.0;
        } else {
            cc[k] = 0.0;
            aa[k] = 1.0;
        }
    }
    waff_suffix<K>(cc, aa, lane);
    // row 0: x0 = (u0 - sum_{j>=1} band_j x_j) / band_0
    double part = 0.0;
    for (int k = 0; k < K; ++k) {
        const int i = K * lane + k;
        if (i >= 1 && i < M) part += band[k] * cc[k];
    }
    const double s = wsum(part, lane);
    const double u0 = wbc(x[0], 0), b0 = wbc(band[0], 0);
    const double x0 = (u0 - s) / b0;
    // forward substitution y_i = (x_i - lo_i y_{i-1}) invd_i
    for (int k = 0; k < K; ++k) {
        const int i = K * lane + k;
        const int n = 2 * i + par;
        if (i == 0) {
            cc[k] = x0;
            aa[k] = 0.0;
        } else if (i < M) {
            const double lo = -(wh_c(n - 2, Nc) * lambda) / (4 * n * (n - 1));
            cc[k] = cc[k] * invd[k];
            aa[k] = -lo * invd[k];
        } else {
            cc[k] = 0.0;
            aa[k] = 1.0;
        }
    }
    waff_prefix<K>(cc, aa, lane);
    for (int k = 0; k < K; ++k) x[k] = (K * lane + k < M) ? cc[k] : 0.0;
}

// Helmholtz solve of a full profile with Dirichlet values ua (y=a), ub (y=b);
// fac points at the packed factors band[N], up[N], invd[N] (contiguous).
template <int K>
__device__ __forceinline__ WV<K> whelm(const WV<K>& f, const double* band, const double* up, const double* invd,
                                       double lambda, double ua, double ub, const WCtx& c) {
    const WV<K> B = wload<K>(band, c), U = wload<K>(up, c), I = wload<K>(invd, c);
    WV<K> x;
    whelm_par<K>(f.e, B.e, U.e, I.e, lambda, (ub + ua) / 2, 0, c.Me, c, x.e);
    whelm_par<K>(f.o, B.o, U.o, I.o, lambda, (ub - ua) / 2, 1, c.Mo, c, x.o);
    return x;
}

// mean value over [a,b] (chebyshev.cpp:505-512): c0 - sum_{n even >= 2} c_n/(n^2-1)
template <int K>
__device__ __forceinline__ double wmean(const WV<K>& v, const WCtx& c) {
    double part = 0.0;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        if (i >= 1 && i < c.Me) part += v.e[k] / (double)(4 * i * i - 1);
    }
    return wbc(v.e[0], 0) - wsum(part, c.lane);
}

// Helmholtz solve with unknown constant forcing mu and mean(u) = umean
// (helmholtz.cpp:158-213); returns mu.
template <int K>
__device__ __forceinline__ double whelm_mean(const WV<K>& f, const double* band, const double* up, const double* invd,
                                             double lambda, double nu, double umean, const WCtx& c, WV<K>& u) {
    const double uamean = wmean<K>(whelm<K>(f, band, up, invd, lambda, 0.0, 0.0, c), c);
    WV<K> r = wzero<K>();
    if (c.lane == 0) r.e[0] = nu;
    const double ucmean = wmean<K>(whelm<K>(r, band, up, invd, lambda, 0.0, 0.0, c), c);
    const double mu = nu * (umean - uamean) / ucmean;
    r = f;
    if (c.lane == 0) r.e[0] += mu;
    u = whelm<K>(r, band, up, invd, lambda, 0.0, 0.0, c);
    return mu;
}

// complex profile <-> LDS row of double2 (re, im) per coefficient n
template <int K>
__device__ __forceinline__ void wload_c(const double2* row, const WCtx& c, WV<K>& re, WV<K>& im) {
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        const double2 a = (i < c.Me) ? row[2 * i] : make_double2(0.0, 0.0);
        const double2 b = (i < c.Mo) ? row[2 * i + 1] : make_double2(0.0, 0.0);
        re.e[k] = a.x;
        im.e[k] = a.y;
        re.o[k] = b.x;
        im.o[k] = b.y;
    }
}

template <int K>
__device__ __forceinline__ void wstore_c(double2* row, const WCtx& c, const WV<K>& re, const WV<K>& im) {
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        if (i < c.Me) row[2 * i] = make_double2(re.e[k], im.e[k]);
        if (i < c.Mo) row[2 * i + 1] = make_double2(re.o[k], im.o[k]);
    }
}

}  // namespace cfg
