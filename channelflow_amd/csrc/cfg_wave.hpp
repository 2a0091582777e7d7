// Wavefront-parallel Chebyshev-tau algebra: one (kx, kz) mode per segment of
// W lanes of a wave64 (W = 16, 32 or 64; 64/W modes per wave). Every shuffle
// is width-W, so scans and reductions never cross a segment.
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

// Global-memory view of a pointer read from a launch table (a generic
// pointer). Loads through it compile to global_load, which only the vector
// memory counter tracks; through the generic pointer they are flat loads,
// which the LDS counter tracks too, so every wait for an LDS access would
// also wait for them.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem_w(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
using gcdouble = const __attribute__((address_space(1))) double*;
// a complex value as a native two-double vector (one 16-byte load or store)
typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int K>
struct WV {
    double e[K], o[K];
};

struct WCtx {
    int N, Me, Mo, Nb, Nc, lane;  // lane: index within the mode's segment
    double L;
    // rows of the Chebyshev-tau operator B of the Helmholtz systems by
    // coefficient n: hb[n], hb[N + n], hb[2N + n] multiply f[n-2], f[n], f[n+2]
    // (zero on rows n < 2 and where B has no entry)
    const double* hb;
};

// DPP move of a double within rows of 16 lanes (a VALU operation, no LDS):
// lanes whose source lies outside the row get 0.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// The same move with 1.0 in the lanes whose source lies outside the row: the
// low word of 1.0 is 0 (zero fill), the high word is kept from a register
// preset to 0x3FF00000.
template <int CTRL>
__device__ __forceinline__ double dpp_row_one(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0x3FF00000, (int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int DPP_ROW_SHL = 0x100, DPP_ROW_SHR = 0x110;      // + shift 1..15
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E;    // quad_perm [1,0,3,2], [2,3,0,1]
constexpr int DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141;

// v of segment lane + d (wdn) / lane - d (wup). A lane whose source is outside
// the segment gets 0 in 16-lane segments, which are DPP rows, and the scans
// there use that 0 (it adds nothing); in wider segments it gets its own v,
// which every caller discards.
template <int W>
__device__ __forceinline__ double wdn(double v, int d) {
    if constexpr (W == 16) {
        switch (d) {
            case 1: return dpp_row<DPP_ROW_SHL + 1>(v);
            case 2: return dpp_row<DPP_ROW_SHL + 2>(v);
            case 4: return dpp_row<DPP_ROW_SHL + 4>(v);
            case 8: return dpp_row<DPP_ROW_SHL + 8>(v);
            default: break;
        }
    }
    return __shfl_down(v, d, W);
}
template <int W>
__device__ __forceinline__ double wup(double v, int d) {
    if constexpr (W == 16) {
        switch (d) {
            case 1: return dpp_row<DPP_ROW_SHR + 1>(v);
            case 2: return dpp_row<DPP_ROW_SHR + 2>(v);
            case 4: return dpp_row<DPP_ROW_SHR + 4>(v);
            case 8: return dpp_row<DPP_ROW_SHR + 8>(v);
            default: break;
        }
    }
    return __shfl_up(v, d, W);
}
// wdn / wup for the multiplier of an affine scan: in 16-lane segments a lane
// whose source is outside the segment gets 1.0, which with the 0 of the
// constant term is the identity map, so those scan steps need no lane
// condition (x + a * 0 = x and a * 1 = a). Wider segments as wdn / wup.
template <int W>
__device__ __forceinline__ double wdn1(double v, int d) {
    if constexpr (W == 16) {
        switch (d) {
            case 1: return dpp_row_one<DPP_ROW_SHL + 1>(v);
            case 2: return dpp_row_one<DPP_ROW_SHL + 2>(v);
            case 4: return dpp_row_one<DPP_ROW_SHL + 4>(v);
            case 8: return dpp_row_one<DPP_ROW_SHL + 8>(v);
            default: break;
        }
    }
    return __shfl_down(v, d, W);
}
template <int W>
__device__ __forceinline__ double wup1(double v, int d) {
    if constexpr (W == 16) {
        switch (d) {
            case 1: return dpp_row_one<DPP_ROW_SHR + 1>(v);
            case 2: return dpp_row_one<DPP_ROW_SHR + 2>(v);
            case 4: return dpp_row_one<DPP_ROW_SHR + 4>(v);
            case 8: return dpp_row_one<DPP_ROW_SHR + 8>(v);
            default: break;
        }
    }
    return __shfl_up(v, d, W);
}
// DPP rows: lanes past a segment's ends read 0 (wdn, wup) or 1 (wdn1, wup1)
template <int W>
constexpr bool wrow() { return W == 16; }
template <int W>
__device__ __forceinline__ double wbc(double v, int src) { return __shfl(v, src, W); }

// sum over the segment, result identical on every lane. W = 16: butterfly over
// DPP partners (quad xor 1, xor 2, half-row mirror, row mirror); each level
// adds a lane's value to its partner's, and a + b = b + a bit for bit, so all
// lanes end with the same sum. Otherwise a tree to lane 0 and a broadcast.
template <int W>
__device__ __forceinline__ double wsum(double v, int lane) {
    if constexpr (W == 16) {
        v += dpp_row<DPP_QUAD_XOR1>(v);
        v += dpp_row<DPP_QUAD_XOR2>(v);
        v += dpp_row<DPP_ROW_HALF_MIRROR>(v);
        v += dpp_row<DPP_ROW_MIRROR>(v);
        return v;
    } else {
        for (int d = W / 2; d >= 1; d >>= 1) {
            const double y = wdn<W>(v, d);
            if (lane + d < W) v += y;
        }
        return wbc<W>(v, 0);
    }
}

template <int K, int W>
__device__ __forceinline__ WV<K> wzero() {
    WV<K> v;
    for (int k = 0; k < K; ++k) v.e[k] = v.o[k] = 0.0;
    return v;
}

// contiguous coefficients c[0..N) (stride 1); P is a pointer to double in
// any address space
template <int K, int W, class P>
__device__ __forceinline__ WV<K> wload(P p, const WCtx& c) {
    WV<K> v;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        v.e[k] = (i < c.Me) ? p[2 * i] : 0.0;
        v.o[k] = (i < c.Mo) ? p[2 * i + 1] : 0.0;
    }
    return v;
}

// strided coefficients c[n] at p[n * s] (a pointer to double in any address
// space)
template <int K, int W, class P>
__device__ __forceinline__ WV<K> wload_s(P p, long s, const WCtx& c) {
    WV<K> v;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        v.e[k] = (i < c.Me) ? p[(long)(2 * i) * s] : 0.0;
        v.o[k] = (i < c.Mo) ? p[(long)(2 * i + 1) * s] : 0.0;
    }
    return v;
}

template <int K, int W, class P>
__device__ __forceinline__ void wstore_s(P p, long s, const WV<K>& v, const WCtx& c) {
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
template <int K, int W>
__device__ __forceinline__ double wpick(const double (&a)[K], int i, const WCtx& c) {
    const int owner = i / K, slot = i - owner * K;
    double x = 0.0;
    for (int k = 0; k < K; ++k) {
        double t = a[k];
        asm volatile("" : "+v"(t));
        x = (k == slot) ? t : x;
    }
    return wbc<W>(x, owner);
}

// coefficients N-1 (even, N odd) and N-2 (odd) of v, broadcast to all lanes
template <int K, int W>
__device__ __forceinline__ double wtop(const WV<K>& v, const WCtx& c) { return wpick<K, W>(v.e, c.Me - 1, c); }
template <int K, int W>
__device__ __forceinline__ double wtop1(const WV<K>& v, const WCtx& c) { return wpick<K, W>(v.o, c.Mo - 1, c); }

// a*x + b*y elementwise
template <int K, int W>
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
template <int K, int W>
__device__ __forceinline__ void wsuffix(const double (&t)[K], double (&incl)[K], double (&excl)[K], int lane) {
    double loc[K];
    double acc = 0.0;
    for (int k = K - 1; k >= 0; --k) {
        acc = t[k] + acc;
        loc[k] = acc;
    }
    double x = loc[0];
    for (int d = 1; d < W; d <<= 1) {
        const double y = wdn<W>(x, d);
        if (wrow<W>() || lane + d < W) x += y;
    }
    double after = wdn<W>(x, 1);
    if (!wrow<W>() && lane == W - 1) after = 0.0;
    for (int k = 0; k < K; ++k) {
        incl[k] = loc[k] + after;
        excl[k] = (k + 1 < K) ? loc[k + 1] + after : after;
    }
}

// d/dy of Chebyshev coefficients on an interval of length L:
// du[2i] = sum_{j>=i} (4/L)(2j+1) c[2j+1],  du[2i+1] = sum_{j>i} (4/L)(2j) c[2j], du[0] halved
template <int K, int W>
__device__ __forceinline__ WV<K> wdiff(const WV<K>& u, const WCtx& c) {
    const double s = 4.0 / c.L;
    double to[K], te[K], inc[K], exc[K];
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        to[k] = (i < c.Mo) ? s * (2 * i + 1) * u.o[k] : 0.0;
        te[k] = (i < c.Me) ? s * (2 * i) * u.e[k] : 0.0;
    }
    WV<K> d;
    wsuffix<K, W>(to, inc, exc, c.lane);
    for (int k = 0; k < K; ++k) d.e[k] = (K * c.lane + k < c.Me) ? inc[k] : 0.0;
    wsuffix<K, W>(te, inc, exc, c.lane);
    for (int k = 0; k < K; ++k) d.o[k] = (K * c.lane + k < c.Mo) ? exc[k] : 0.0;
    if (c.lane == 0) d.e[0] *= 0.5;
    return d;
}

// values at y=b (sum of coefficients) and y=a (alternating sum)
template <int K, int W>
__device__ __forceinline__ void wends(const WV<K>& v, const WCtx& c, double& at_b, double& at_a) {
    double se = 0.0, so = 0.0;
    for (int k = 0; k < K; ++k) {
        se += v.e[k];
        so += v.o[k];
    }
    se = wsum<W>(se, c.lane);
    so = wsum<W>(so, c.lane);
    at_b = se + so;
    at_a = se - so;
}

// Affine suffix scan: x_i = cc_i + aa_i x_{i+1}; maps past the chain end are
// the identity (cc = 0, aa = 1) and the last real map is constant (aa = 0).
// On return cc[k] holds x_i.
template <int K, int W>
__device__ __forceinline__ void waff_suffix(double (&cc)[K], const double (&aa)[K], int lane) {
    double Fc[K], Fa[K];
    double xc = 0.0, xa = 1.0;
    for (int k = K - 1; k >= 0; --k) {
        xc = cc[k] + aa[k] * xc;
        xa = aa[k] * xa;
        Fc[k] = xc;
        Fa[k] = xa;
    }
    for (int d = 1; d < W; d <<= 1) {
        const double yc = wdn<W>(xc, d), ya = wdn1<W>(xa, d);
        if (wrow<W>() || lane + d < W) {
            xc = xc + xa * yc;
            xa = xa * ya;
        }
    }
    double ec = wdn<W>(xc, 1);
    if (!wrow<W>() && lane == W - 1) ec = 0.0;
    for (int k = 0; k < K; ++k) cc[k] = Fc[k] + Fa[k] * ec;
}

// Affine prefix scan: y_i = cc_i + aa_i y_{i-1}; the first map is constant.
template <int K, int W>
__device__ __forceinline__ void waff_prefix(double (&cc)[K], const double (&aa)[K], int lane) {
    double Gc[K], Ga[K];
    double xc = 0.0, xa = 1.0;
    for (int k = 0; k < K; ++k) {
        xc = cc[k] + aa[k] * xc;
        xa = aa[k] * xa;
        Gc[k] = xc;
        Ga[k] = xa;
    }
    for (int d = 1; d < W; d <<= 1) {
        const double yc = wup<W>(xc, d), ya = wup1<W>(xa, d);
        if (wrow<W>() || lane >= d) {
            xc = xc + xa * yc;
            xa = xa * ya;
        }
    }
    double ec = wup<W>(xc, 1);
    if (!wrow<W>() && lane == 0) ec = 0.0;
    for (int k = 0; k < K; ++k) cc[k] = Gc[k] + Ga[k] * ec;
}

// ---- several right-hand sides per scan ----------------------------------------
// The Chebyshev-tau operators are real, so the real and imaginary parts of a
// mode (and u and w, which share the velocity operator) are solved together:
// one affine scan carries NR constant terms with a shared multiplier, which
// keeps the arithmetic of each right-hand side exactly that of the one-RHS
// routines above while the dependent shuffle chains are shared.

template <int K, int W, int NR>
__device__ __forceinline__ void waff_suffix_n(double (&cc)[NR][K], const double (&aa)[K], int lane) {
    double Fc[NR][K], Fa[K];
    double xc[NR], xa = 1.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) xc[r] = 0.0;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            xc[r] = cc[r][k] + aa[k] * xc[r];
            Fc[r][k] = xc[r];
        }
        xa = aa[k] * xa;
        Fa[k] = xa;
    }
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        double yc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) yc[r] = wdn<W>(xc[r], d);
        const double ya = wdn1<W>(xa, d);
        if (wrow<W>() || lane + d < W) {
#pragma unroll
            for (int r = 0; r < NR; ++r) xc[r] = xc[r] + xa * yc[r];
            xa = xa * ya;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double ec = wdn<W>(xc[r], 1);
        if (!wrow<W>() && lane == W - 1) ec = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) cc[r][k] = Fc[r][k] + Fa[k] * ec;
    }
}

template <int K, int W, int NR>
__device__ __forceinline__ void waff_prefix_n(double (&cc)[NR][K], const double (&aa)[K], int lane) {
    double Gc[NR][K], Ga[K];
    double xc[NR], xa = 1.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) xc[r] = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            xc[r] = cc[r][k] + aa[k] * xc[r];
            Gc[r][k] = xc[r];
        }
        xa = aa[k] * xa;
        Ga[k] = xa;
    }
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        double yc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) yc[r] = wup<W>(xc[r], d);
        const double ya = wup1<W>(xa, d);
        if (wrow<W>() || lane >= d) {
#pragma unroll
            for (int r = 0; r < NR; ++r) xc[r] = xc[r] + xa * yc[r];
            xa = xa * ya;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double ec = wup<W>(xc[r], 1);
        if (!wrow<W>() && lane == 0) ec = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) cc[r][k] = Gc[r][k] + Ga[k] * ec;
    }
}

// sums over the wave of NR values, identical on every lane (as wsum)
template <int W, int NR>
__device__ __forceinline__ void wsum_n(double (&v)[NR], int lane) {
    if constexpr (W == 16) {
#pragma unroll
        for (int r = 0; r < NR; ++r) v[r] = wsum<W>(v[r], lane);
        return;
    }
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) {
        double y[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) y[r] = wdn<W>(v[r], d);
        if (lane + d < W) {
#pragma unroll
            for (int r = 0; r < NR; ++r) v[r] += y[r];
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wbc<W>(v[r], 0);
}

// One parity system of nu u'' - lambda u = f for NR right-hand sides f[r]
// with u[0] set to bc[r]: x = A^{-1} (B f), B from the rows bl, bd, bu of this
// parity and A from its UL factors band / up / invd (helmholtz.cpp:79-95 with
// bandedtridiag.cpp:232-277). One affine scan carries all NR constant terms
// with a shared multiplier, so each right-hand side gets the arithmetic it
// would get alone.
template <int K, int W, int NR>
__device__ __forceinline__ void whelm_par_n(const double (&f)[NR][K], const double (&band)[K], const double (&up)[K],
                                            const double (&invd)[K], const double (&bl)[K], const double (&bd)[K],
                                            const double (&bu)[K], double lambda, const double (&bc)[NR], int M,
                                            const WCtx& c, double (&x)[NR][K]) {
    const int lane = c.lane;
    double cc[NR][K], aa[K];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double fm[K], fp[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            fm[k] = (k > 0) ? f[r][k - 1] : 0.0;
            fp[k] = (k + 1 < K) ? f[r][k + 1] : 0.0;
        }
        const double lastp = wup<W>(f[r][K - 1], 1);
        const double firstn = wdn<W>(f[r][0], 1);
        if (wrow<W>() || lane > 0) fm[0] = lastp;
        fp[K - 1] = (wrow<W>() || lane < W - 1) ? firstn : 0.0;
        // u = B f: zero past the system (zero rows), bc on row 0
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double u = bl[k] * fm[k] + bd[k] * f[r][k] + bu[k] * fp[k];
            x[r][k] = u;
            cc[r][k] = u;
        }
        if (lane == 0) {
            x[r][0] = bc[r];
            cc[r][0] = 0.0;  // row 0 is not a backward-substitution map
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = K * lane + k;
        aa[k] = (i >= 1 && i < M - 1) ? -up[k] : ((i == M - 1 && i >= 1) ? 0.0 : 1.0);
    }
    waff_suffix_n<K, W, NR>(cc, aa, lane);
    // row 0: x0 = (u0 - sum_{j>=1} band_j x_j) / band_0 (band is zero past the system)
    double part[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        part[r] = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k == 0 && lane == 0) continue;
            part[r] += band[k] * cc[r][k];
        }
    }
    wsum_n<W, NR>(part, lane);
    // x0 is used by lane 0 only (row i = 0), which holds u0 = x[r][0] and band_0
    double x0[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) x0[r] = (x[r][0] - part[r]) / band[0];
    // forward substitution y_i = (x_i - lo_i y_{i-1}) invd_i with lo_i = -lambda bl_i
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = K * lane + k;
        const bool in = i < M;
        aa[k] = in ? (lambda * bl[k]) * invd[k] : 1.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) cc[r][k] = in ? cc[r][k] * invd[k] : 0.0;
    }
    if (lane == 0) {
        aa[0] = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) cc[r][0] = x0[r];
    }
    waff_prefix_n<K, W, NR>(cc, aa, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) x[r][k] = (K * lane + k < M) ? cc[r][k] : 0.0;
}

// Helmholtz solve of NR profiles sharing one operator with Dirichlet values
// ua[r] (y=a), ub[r] (y=b); band, up, invd: the operator's UL factors, N each.
template <int K, int W, int NR, class P>
__device__ __forceinline__ void whelm_n(const WV<K> (&f)[NR], P band, P up, P invd, double lambda,
                                        const double (&ua)[NR], const double (&ub)[NR], const WCtx& c,
                                        WV<K> (&x)[NR]) {
    const WV<K> B = wload<K, W>(band, c), U = wload<K, W>(up, c), I = wload<K, W>(invd, c);
    const WV<K> HL = wload<K, W>(c.hb, c), HD = wload<K, W>(c.hb + c.N, c), HU = wload<K, W>(c.hb + 2 * c.N, c);
    double fe[NR][K], fo[NR][K], xe[NR][K], xo[NR][K], be[NR], bo[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        be[r] = (ub[r] + ua[r]) / 2;
        bo[r] = (ub[r] - ua[r]) / 2;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            fe[r][k] = f[r].e[k];
            fo[r][k] = f[r].o[k];
        }
    }
    whelm_par_n<K, W, NR>(fe, B.e, U.e, I.e, HL.e, HD.e, HU.e, lambda, be, c.Me, c, xe);
    whelm_par_n<K, W, NR>(fo, B.o, U.o, I.o, HL.o, HD.o, HU.o, lambda, bo, c.Mo, c, xo);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            x[r].e[k] = xe[r][k];
            x[r].o[k] = xo[r][k];
        }
}

// one profile
template <int K, int W, class P>
__device__ __forceinline__ WV<K> whelm(const WV<K>& f, P band, P up, P invd, double lambda, double ua, double ub,
                                       const WCtx& c) {
    const WV<K> f1[1] = {f};
    const double a1[1] = {ua}, b1[1] = {ub};
    WV<K> x1[1];
    whelm_n<K, W, 1>(f1, band, up, invd, lambda, a1, b1, c, x1);
    return x1[0];
}

// mean value over [a,b] (chebyshev.cpp:505-512): c0 - sum_{n even >= 2} c_n/(n^2-1)
template <int K, int W>
__device__ __forceinline__ double wmean(const WV<K>& v, const WCtx& c) {
    double part = 0.0;
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        if (i >= 1 && i < c.Me) part += v.e[k] / (double)(4 * i * i - 1);
    }
    return wbc<W>(v.e[0], 0) - wsum<W>(part, c.lane);
}

// Helmholtz solve with unknown constant forcing mu and mean(u) = umean
// (helmholtz.cpp:158-213); returns mu.
template <int K, int W, class P>
__device__ __forceinline__ double whelm_mean(const WV<K>& f, P band, P up, P invd, double lambda, double nu,
                                             double umean, const WCtx& c, WV<K>& u) {
    const double uamean = wmean<K, W>(whelm<K, W>(f, band, up, invd, lambda, 0.0, 0.0, c), c);
    WV<K> r = wzero<K, W>();
    if (c.lane == 0) r.e[0] = nu;
    const double ucmean = wmean<K, W>(whelm<K, W>(r, band, up, invd, lambda, 0.0, 0.0, c), c);
    const double mu = nu * (umean - uamean) / ucmean;
    r = f;
    if (c.lane == 0) r.e[0] += mu;
    u = whelm<K, W>(r, band, up, invd, lambda, 0.0, 0.0, c);
    return mu;
}


// suffix sums of NR arrays at once (as wsuffix)
template <int K, int W, int NR>
__device__ __forceinline__ void wsuffix_n(const double (&t)[NR][K], double (&incl)[NR][K], double (&excl)[NR][K],
                                          int lane) {
    double loc[NR][K], x[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            acc = t[r][k] + acc;
            loc[r][k] = acc;
        }
        x[r] = loc[r][0];
    }
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        double y[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) y[r] = wdn<W>(x[r], d);
        if (wrow<W>() || lane + d < W) {
#pragma unroll
            for (int r = 0; r < NR; ++r) x[r] += y[r];
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double after = wdn<W>(x[r], 1);
        if (!wrow<W>() && lane == W - 1) after = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            incl[r][k] = loc[r][k] + after;
            excl[r][k] = (k + 1 < K) ? loc[r][k + 1] + after : after;
        }
    }
}

// d/dy of NR profiles (as wdiff)
template <int K, int W, int NR>
__device__ __forceinline__ void wdiff_n(const WV<K> (&u)[NR], const WCtx& c, WV<K> (&d)[NR]) {
    const double s = 4.0 / c.L;
    double to[NR][K], te[NR][K], inc[NR][K], exc[NR][K];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = K * c.lane + k;
            to[r][k] = (i < c.Mo) ? s * (2 * i + 1) * u[r].o[k] : 0.0;
            te[r][k] = (i < c.Me) ? s * (2 * i) * u[r].e[k] : 0.0;
        }
    wsuffix_n<K, W, NR>(to, inc, exc, c.lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) d[r].e[k] = (K * c.lane + k < c.Me) ? inc[r][k] : 0.0;
    wsuffix_n<K, W, NR>(te, inc, exc, c.lane);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
#pragma unroll
        for (int k = 0; k < K; ++k) d[r].o[k] = (K * c.lane + k < c.Mo) ? exc[r][k] : 0.0;
        if (c.lane == 0) d[r].e[0] *= 0.5;
    }
}

// values at y=b and y=a of NR profiles (as wends)
template <int K, int W, int NR>
__device__ __forceinline__ void wends_n(const WV<K> (&v)[NR], const WCtx& c, double (&at_b)[NR], double (&at_a)[NR]) {
    double s2[2 * NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double se = 0.0, so = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            se += v[r].e[k];
            so += v[r].o[k];
        }
        s2[2 * r] = se;
        s2[2 * r + 1] = so;
    }
    wsum_n<W, 2 * NR>(s2, c.lane);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        at_b[r] = s2[2 * r] + s2[2 * r + 1];
        at_a[r] = s2[2 * r] - s2[2 * r + 1];
    }
}

// complex profile <-> LDS row of double2 (re, im) per coefficient n
template <int K, int W>
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

template <int K, int W>
__device__ __forceinline__ void wstore_c(double2* row, const WCtx& c, const WV<K>& re, const WV<K>& im) {
    for (int k = 0; k < K; ++k) {
        const int i = K * c.lane + k;
        if (i < c.Me) row[2 * i] = make_double2(re.e[k], im.e[k]);
        if (i < c.Mo) row[2 * i + 1] = make_double2(re.o[k], im.o[k]);
    }
}

}  // namespace cfg
