// HIP kernels of the MI355X DNS time step (gfx950, FP64 throughout).
//
// Nonlinear term pipeline for one field (reference: NSE::nonlinear,
// channelflow/nse.cpp:383-391 -> diffops.cpp rotationalNL/convectionNL/
// divergenceNL/skewsymmetricNL):
//   k_yinv_t Chebyshev coefficients -> values at the Chebyshev points y_n,
//            with d/dy folded into the operator (T, T*D) and the base flow
//            U(y), W(y), -Vsuck added to the kx=kz=0 mode;
//   k_xwinv  inverse x-FFT per (field, y, kz) line, optional d/dx (k_xinv:
//            LDS FFTs for the lengths the wavefront FFT does not cover);
//   k_zwave  per (y, x) line: c2r z-FFTs, pointwise products, r2c z-FFTs,
//            a wavefront per group of lines (k_zpass: LDS FFTs, block-wide,
//            for the lengths the wavefront FFT does not cover);
//   k_xwfwd  forward x-FFT, live kx kept, 1/(Nx Nz) scaling (k_xfwd);
//   k_yfwd_t values -> Chebyshev coefficients, with the spectral divergence
//            of u_i u_j folded in for the divergence/skew-symmetric forms.
// Only non-aliased modes |kx| <= kxd, kz <= kzd are ever stored, so the
// reference's zeroPaddedModes() (flowfield.cpp:2235) is implicit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfg_kernels.hpp"
#include "cfg_math.hpp"
#include "cfg_wave.hpp"

namespace cfg {

constexpr double PI = 3.14159265358979323846;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 imul(double s, double2 a) { return make_double2(-(s * a.y), s * a.x); }  // (i s) a

__device__ __forceinline__ int kx_of(const Geom& g, int ikx) { return ikx <= g.kxd ? ikx : ikx - g.nkx; }

// LDS layout of FFT lines: element i of a line sits at pidx(i) = i + i/8 (one
// pad slot per 8 complex values spreads the power-of-two strides of the
// Stockham passes over distinct banks), and consecutive lines are lpitch(N) =
// N + N/8 + 1 slots apart (odd, so walks across lines are conflict-free too).
__host__ __device__ constexpr int pidx(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int lpitch(int N) { return N + N / 8 + 1; }

// ---------------------------------------------------------------------------
// LDS Stockham FFTs over nlines contiguous lines of length N. sign -1 forward
// (exp(-i...)), +1 backward; unnormalised. tw is the table exp(-2 pi i j/N),
// j < N, staged in LDS. The result lands in a or b (returned). All threads of
// the block must call them.
// ---------------------------------------------------------------------------
template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2 (&v)[4]) {
    double2 o[4];
    if constexpr (R == 2) {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        const double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const double2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        const double2 jd = (SIGN < 0) ? make_double2(d13.y, -d13.x) : make_double2(-d13.y, d13.x);
        o[0] = cadd(s02, s13);
        o[1] = cadd(d02, jd);
        o[2] = csub(s02, s13);
        o[3] = csub(d02, jd);
    } else {
        const double s3 = (SIGN < 0) ? -0.86602540378443864676 : 0.86602540378443864676;
        const double2 s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
        const double2 m = make_double2(v[0].x - 0.5 * s12.x, v[0].y - 0.5 * s12.y);
        o[0] = cadd(v[0], s12);
        o[1] = make_double2(fma(-s3, d12.y, m.x), fma(s3, d12.x, m.y));  // m + q, q = i s3 d12
        o[2] = make_double2(fma(s3, d12.y, m.x), fma(-s3, d12.x, m.y));  // m - q
    }
    for (int r = 0; r < R; ++r) v[r] = o[r];
}

// radix 8, decimation in frequency: sums and differences of v_k, v_{k+4}, the
// differences turned by w^k (w = exp(SIGN 2 pi i/8)), then two radix-4 DFTs
// give the even and the odd outputs. The 1/sqrt(2) of w and w^3 is applied in
// the odd radix-4's first sums, as fused multiply-adds: with p1 = w b1 sqrt2
// and p3 = w^3 b3 sqrt2, b1 + b3 = h (p1 + p3) and b1 - b3 = h (p1 - p3).
template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2 (&v)[8]) {
    static_assert(R == 8, "radix-8 overload");
    constexpr double h = 0.70710678118654752440;  // 1/sqrt(2)
    double2 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = cadd(v[k], v[k + 4]);
        b[k] = csub(v[k], v[k + 4]);
    }
    double2 p1, p3;
    if constexpr (SIGN < 0) {  // w = (1 - i)/sqrt2, w^2 = -i, w^3 = -(1 + i)/sqrt2
        p1 = make_double2(b[1].x + b[1].y, b[1].y - b[1].x);
        p3 = make_double2(b[3].y - b[3].x, -(b[3].x + b[3].y));
        b[2] = make_double2(b[2].y, -b[2].x);
    } else {  // w = (1 + i)/sqrt2, w^2 = i, w^3 = (-1 + i)/sqrt2
        p1 = make_double2(b[1].x - b[1].y, b[1].x + b[1].y);
        p3 = make_double2(-(b[3].x + b[3].y), b[3].x - b[3].y);
        b[2] = make_double2(-b[2].y, b[2].x);
    }
    dft_small<4, SIGN>(a);
    // radix 4 of (b0, h p1, b2, h p3), as in dft_small<4>
    const double2 s02 = cadd(b[0], b[2]), d02 = csub(b[0], b[2]);
    const double2 s13 = cadd(p1, p3), d13 = csub(p1, p3);                                // over h
    const double2 jd = (SIGN < 0) ? make_double2(d13.y, -d13.x) : make_double2(-d13.y, d13.x);  // over h
    b[0] = make_double2(fma(h, s13.x, s02.x), fma(h, s13.y, s02.y));
    b[1] = make_double2(fma(h, jd.x, d02.x), fma(h, jd.y, d02.y));
    b[2] = make_double2(fma(-h, s13.x, s02.x), fma(-h, s13.y, s02.y));
    b[3] = make_double2(fma(-h, jd.x, d02.x), fma(-h, jd.y, d02.y));
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = a[m];
        v[2 * m + 1] = b[m];
    }
}

template <int R, int SIGN>
__device__ __forceinline__ void stockham_bfly(const double2* __restrict__ s, double2* __restrict__ d, const double2* tw,
                                              int j, int k, int NR, int NS, int tstep) {
    double2 v[4];
    for (int r = 0; r < R; ++r) v[r] = s[pidx(j + r * NR)];
    if (k != 0) {
        for (int r = 1; r < R; ++r) {
            double2 t = tw[r * k * tstep];
            if (SIGN > 0) t.y = -t.y;
            v[r] = cmul(v[r], t);
        }
    }
    dft_small<R, SIGN>(v);
    const int o = (j / NS) * NS * R + k;
    for (int r = 0; r < R; ++r) d[pidx(o + r * NS)] = v[r];
}

// runtime length (plan in P)
template <int SIGN>
__device__ double2* fft_lds(double2* a, double2* b, const FFTPlan& P, const double2* tw, int nlines) {
    double2* src = a;
    double2* dst = b;
    const int N = P.N;
    int Ns = 1;
    for (int p = 0; p < P.npass; ++p) {
        const int R = P.radix[p];
        const int NR = N / R;
        const int tstep = N / (Ns * R);
        for (int w = threadIdx.x; w < nlines * NR; w += blockDim.x) {
            const int line = w / NR;
            const int j = w - line * NR;
            const int k = j % Ns;
            const int lo = line * lpitch(N);
            if (R == 4) stockham_bfly<4, SIGN>(src + lo, dst + lo, tw, j, k, NR, Ns, tstep);
            else if (R == 2) stockham_bfly<2, SIGN>(src + lo, dst + lo, tw, j, k, NR, Ns, tstep);
            else stockham_bfly<3, SIGN>(src + lo, dst + lo, tw, j, k, NR, Ns, tstep);
        }
        __syncthreads();
        double2* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

// compile-time length: every divisor is a constant. Radices 4, 2, 3 in that
// order; with r8, radix 8 first where that takes fewer passes (ties: 4). The
// z-pass gains from the fewer passes; the x-FFTs, with fewer lines per block,
// lose more to the smaller number of butterflies per pass.
constexpr int fft_passes_from(int n, bool eight) {
    int c = 0;
    while (eight && n % 8 == 0) n /= 8, ++c;
    while (n % 4 == 0) n /= 4, ++c;
    while (n % 2 == 0) n /= 2, ++c;
    while (n % 3 == 0) n /= 3, ++c;
    return c;
}
constexpr int fft_radix(int n, int p, bool r8) {
    const bool eight = r8 && fft_passes_from(n, true) < fft_passes_from(n, false);
    int c = 0;
    while (eight && n % 8 == 0) {
        if (c == p) return 8;
        n /= 8;
        ++c;
    }
    while (n % 4 == 0) {
        if (c == p) return 4;
        n /= 4;
        ++c;
    }
    while (n % 2 == 0) {
        if (c == p) return 2;
        n /= 2;
        ++c;
    }
    while (n % 3 == 0) {
        if (c == p) return 3;
        n /= 3;
        ++c;
    }
    return 0;
}
constexpr int fft_npass(int n, bool r8) {
    int c = 0;
    while (fft_radix(n, c, r8) != 0) ++c;
    return c;
}
constexpr int fft_ns(int n, int p, bool r8) {
    int s = 1;
    for (int q = 0; q < p; ++q) s *= fft_radix(n, q, r8);
    return s;
}

// Threads per block of every kernel that runs the compile-time FFTs.
constexpr int FFT_BLOCK = 256;

// One Stockham pass per recursion level. A thread keeps the same butterfly
// index j in every line it visits (FFT_BLOCK / NR lines per sweep), so its
// padded source and destination offsets and its twiddles are computed once
// per pass; a twiddle of 1 (k = 0) multiplies exactly.
template <int N, int SIGN, int P, bool R8>
__device__ __forceinline__ double2* fft_ct(double2* a, double2* b, const double2* tw, int nlines) {
    if constexpr (P == fft_npass(N, R8)) {
        return a;
    } else {
        constexpr int R = fft_radix(N, P, R8), NR = N / R, NS = fft_ns(N, P, R8), TSTEP = N / (NS * R);
        constexpr int LPI = FFT_BLOCK / NR, LP = lpitch(N);
        static_assert(LPI >= 1, "FFT line longer than the block");
        const int j = threadIdx.x % NR, l0 = threadIdx.x / NR;
        if (l0 < LPI) {
            const int k = j % NS;
            const int o = (j / NS) * NS * R + k;
            int so[R], dof[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                so[r] = pidx(j + r * NR);
                dof[r] = pidx(o + r * NS);
            }
            double2 t[R];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    t[r] = tw[r * k * TSTEP];
                    if (SIGN > 0) t[r].y = -t[r].y;
                }
            }
            for (int line = l0; line < nlines; line += LPI) {
                const double2* __restrict__ src = a + line * LP;
                double2* __restrict__ dst = b + line * LP;
                double2 v[R == 8 ? 8 : 4];
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = src[so[r]];
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], t[r]);
                }
                dft_small<R, SIGN>(v);
#pragma unroll
                for (int r = 0; r < R; ++r) dst[dof[r]] = v[r];
            }
        }
        __syncthreads();
        return fft_ct<N, SIGN, P + 1, R8>(b, a, tw, nlines);
    }
}

// N == 0 selects the runtime-length path
template <int N, int SIGN, bool R8 = false>
__device__ __forceinline__ double2* fft_any(double2* a, double2* b, const FFTPlan& P, const double2* tw, int nlines) {
    if constexpr (N == 0) return fft_lds<SIGN>(a, b, P, tw, nlines);
    else return fft_ct<N, SIGN, 0, R8>(a, b, tw, nlines);
}

// stage the twiddle table of P into LDS
__device__ __forceinline__ void stage_tw(double2* dst, const FFTPlan& P) {
    for (int i = threadIdx.x; i < P.N; i += blockDim.x) dst[i] = P.tw[i];
}

// ===========================================================================
// Tau solver kernels: one thread per live mode (tausolver.cpp, nse.cpp:479)
// ===========================================================================
__device__ TauView make_view(const Geom& g, const TauSet& s, int m, double nu, bool taucorr) {
    const int ikx = m % g.nkx, ikz = m / g.nkx;
    const int kx = kx_of(g, ikx), kz = ikz;
    const double pi2 = PI * PI;
    const double kappa2 = 4 * pi2 * ((kx / g.Lx) * (kx / g.Lx) + (kz / g.Lz) * (kz / g.Lz));
    const double c = 4.0 * pi2 * nu;
    const double lambda = s.lambda_t + c * ((kx / g.Lx) * (kx / g.Lx) + (kz / g.Lz) * (kz / g.Lz));
    double* cm = s.coef + (long)m * TAU_NA * g.Ny;
    auto A = [&](int k) { return SV{cm + (long)k * g.Ny, 1}; };
    TauView t;
    t.ph = HelmholtzView{A(0), A(1), A(2), kappa2, g.Ny};
    t.vh = HelmholtzView{A(3), A(4), A(5), lambda, g.Ny};
    t.Pp = A(6);
    t.Pm = A(7);
    t.vp = A(8);
    t.vm = A(9);
    t.P0 = A(10);
    t.v0 = A(11);
    t.scal = s.scal + (long)m * 6;
    t.sstride = 1;
    t.nu = nu;
    t.a = g.a;
    t.b = g.b;
    t.kx2pi = 2 * PI * kx / g.Lx;
    t.kz2pi = 2 * PI * kz / g.Lz;
    t.zero_mode = (kx == 0 && kz == 0);
    t.taucorrection = taucorr;
    return t;
}

// TauSolver constructor per mode (tausolver.cpp:81-176), T modes per block,
// one thread each. ALL: the mode's whole working set in LDS, its 12 tau
// profiles (array a at element (a*Ny + n)*T + t), 6 scalars and 2 scratch
// profiles, copied to the mode-major tau set with contiguous writes at the end.
// Otherwise (tall grids, where a block of 64 KB would hold fewer than 8 such
// working sets) the profiles and scalars are built in place in the tau set and
// only the 2 scratch profiles are in LDS. The arithmetic is tau_setup's either
// way (the sequential set-up, bit for bit the oracle's); only the memory the
// recurrences store to and reload from moves.
template <bool ALL>
__global__ void k_tau_setup(Geom g, TauSet s, double nu, int taucorr) {
    extern __shared__ double sl[];
    const int T = blockDim.x, t = threadIdx.x;
    const int m0 = blockIdx.x * T, m = m0 + t;
    const int Ny = g.Ny;
    const long arr = (long)Ny * T;  // one profile of the block's modes
    if (m < g.M) {
        TauView v = make_view(g, s, m, nu, taucorr != 0);
        auto A = [&](int a) { return SV{sl + a * arr + t, T}; };
        if constexpr (ALL) {
            v.ph.band = A(0);
            v.ph.up = A(1);
            v.ph.invd = A(2);
            v.vh.band = A(3);
            v.vh.up = A(4);
            v.vh.invd = A(5);
            v.Pp = A(6);
            v.Pm = A(7);
            v.vp = A(8);
            v.vm = A(9);
            v.P0 = A(10);
            v.v0 = A(11);
            v.scal = sl + 14 * arr + t;
            v.sstride = T;
            tau_setup(v, A(12), A(13));
        } else {
            tau_setup(v, A(0), A(1));
        }
    }
    if constexpr (ALL) {
        __syncthreads();
        const int nm = min(T, g.M - m0);
        __attribute__((address_space(1))) double* coef = gmem_w(s.coef) + (long)m0 * TAU_NA * Ny;
        for (int i = t; i < nm * TAU_NA * Ny; i += T) {
            const int mm = i / (TAU_NA * Ny), r = i - mm * (TAU_NA * Ny), a = r / Ny, n = r - a * Ny;
            coef[i] = sl[a * arr + (long)n * T + mm];
        }
        __attribute__((address_space(1))) double* scal = gmem_w(s.scal) + (long)m0 * 6;
        for (int i = t; i < nm * 6; i += T) scal[i] = sl[14 * arr + (long)(i % 6) * T + i / 6];
    }
}

// modes per block of the set-up within 64 KB of LDS, at most 64: whole working
// sets (14 profiles and 6 scalars each) where 8 or more fit, else scratch
// profiles only (2 each)
static int tau_setup_modes(int Ny, bool* all) {
    const long whole = (14L * Ny + 6) * (long)sizeof(double), scratch = 2L * Ny * (long)sizeof(double);
    const long T = (64L * 1024) / whole;
    *all = T >= 8;
    return (int)std::max(1L, std::min(64L, (64L * 1024) / (*all ? whole : scratch)));
}

void launch_tau_setup(const Geom& g, const TauSet& s, double nu, bool taucorr, hipStream_t st) {
    bool all = false;
    const int T = tau_setup_modes(g.Ny, &all);
    const dim3 grid((g.M + T - 1) / T);
    if (all) {
        const size_t lds = (size_t)(14L * g.Ny + 6) * T * sizeof(double);
        k_tau_setup<true><<<grid, T, lds, st>>>(g, s, nu, taucorr ? 1 : 0);
    } else {
        const size_t lds = (size_t)2 * g.Ny * T * sizeof(double);
        k_tau_setup<false><<<grid, T, lds, st>>>(g, s, nu, taucorr ? 1 : 0);
    }
}

__device__ __forceinline__ CV cview(double* base, const Geom& g, int comp, int m) {
    const long S2 = 2L * g.M;
    double* p = base + 2L * ((long)comp * g.Ny * g.M + m);
    return CV{SV{p, S2}, SV{p + 1, S2}};
}

// Tau solve with one W-lane segment of a wavefront per live mode (NSE::solve, nse.cpp:479-575;
// TauSolver::solve, tausolver.cpp:347-450, with solve_P_and_v :193-251).

// solve_P_and_v for the real and imaginary parts at once (NR = 2 right-hand
// sides per scan); per part the same arithmetic as the one-RHS routines
// (tausolver.cpp:193-251).
// Ry is read again from its LDS row (re, im per coefficient) where it is
// needed, instead of being held in registers through the pressure solve.
template <int K, int W>
__device__ __forceinline__ void tau_P_and_v_wave2(const WCtx& c, gcdouble cm, gcdouble sc, double kappa2,
                                                  double lambda, double nu, bool taucorr, const WV<K> (&r)[2],
                                                  const double2* Ryrow, WV<K> (&P)[2], WV<K> (&v)[2]) {
    const long Ny = c.N;
    const double z2[2] = {0.0, 0.0};
    whelm_n<K, W, 2>(r, cm + 0 * Ny, cm + 1 * Ny, cm + 2 * Ny, kappa2, z2, z2, c, P);
    WV<K> tmp[2];
    wdiff_n<K, W, 2>(P, c, tmp);
    double rtop[2], rtop1[2];  // Ry's top coefficients, for the tau correction
    {
        WV<K> Ry[2];
        wload_c<K, W>(Ryrow, c, Ry[0], Ry[1]);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                tmp[p].e[k] -= Ry[p].e[k];
                tmp[p].o[k] -= Ry[p].o[k];
            }
            rtop[p] = wtop<K, W>(Ry[p], c);
            rtop1[p] = wtop1<K, W>(Ry[p], c);
        }
    }
    whelm_n<K, W, 2>(tmp, cm + 3 * Ny, cm + 4 * Ny, cm + 5 * Ny, lambda, z2, z2, c, v);
    // influence-matrix correction (tausolver.cpp:178-191)
    double dvb[2], dva[2];
    wdiff_n<K, W, 2>(v, c, tmp);
    wends_n<K, W, 2>(tmp, c, dvb, dva);
    {
        const WV<K> Pp = wload<K, W>(cm + 6 * Ny, c), Pm = wload<K, W>(cm + 7 * Ny, c);
        const WV<K> vp = wload<K, W>(cm + 8 * Ny, c), vm = wload<K, W>(cm + 9 * Ny, c);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double dplus = -sc[0] * dvb[p] - sc[1] * dva[p];
            const double dminus = -sc[2] * dvb[p] - sc[3] * dva[p];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                P[p].e[k] += dplus * Pp.e[k] + dminus * Pm.e[k];
                P[p].o[k] += dplus * Pp.o[k] + dminus * Pm.o[k];
                v[p].e[k] += dplus * vp.e[k] + dminus * vm.e[k];
                v[p].o[k] += dplus * vp.o[k] + dminus * vm.o[k];
            }
        }
    }
    if (!taucorr) return;
    const int Nb = c.Nb;
    const WV<K> P0 = wload<K, W>(cm + 10 * Ny, c), v0 = wload<K, W>(cm + 11 * Ny, c);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double sigma1_Nb = lambda * wtop<K, W>(v[p], c) - nu * 0.0 - rtop[p];
        double sigma1_Nb1 = lambda * wtop1<K, W>(v[p], c) - nu * 0.0 - rtop1[p];
        sigma1_Nb1 += (4.0 / c.L) * Nb * wtop<K, W>(P[p], c);
        const double sigmaNb = sigma1_Nb / (1.0 - sc[5]);
        const double sigmaNb1 = sigma1_Nb1 / (1.0 - sc[4]);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            P[p].e[k] += sigmaNb1 * P0.e[k];
            P[p].o[k] += sigmaNb * P0.o[k];
            v[p].e[k] += sigmaNb * v0.e[k];
            v[p].o[k] += sigmaNb1 * v0.o[k];
        }
    }
}

// The right-hand-side rows [c][n][m0..m0+MPB) of a block: sum_j coef_j in_j
// (the multistep / Runge-Kutta combination, dnsalgo.cpp:218-224, 668-673),
// stored to LDS as [mode][c][n]. The term count is a template parameter and
// UNR elements are loaded per pass, so the NT x UNR loads of a pass are all in
// flight before the first is summed.
template <int NT, int MPB>
__device__ __forceinline__ void tau_stage_rhs(const TauMember& me, const Geom& g, int m0, int nthr, double2* lds) {
    constexpr int UNR = NT <= 2 ? 8 : (NT <= 4 ? 4 : 2);  // 16 loads in flight at most
    const __attribute__((address_space(1))) dvec2* in[NT];
    double coef[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        in[j] = gmem(reinterpret_cast<const dvec2*>(me.in[j]));
        coef[j] = me.coef[j];
    }
    const int Ny = g.Ny;
    const int total = 3 * Ny * MPB;
    for (int base = threadIdx.x; base < total; base += UNR * nthr) {
        dvec2 v[UNR][NT];
        int slot[UNR];
#pragma unroll
        for (int r = 0; r < UNR; ++r) {
            const int idx = base + r * nthr;
            const int ww = idx % MPB, rest = idx / MPB, n = rest % Ny, comp = rest / Ny;
            const int mm = m0 + ww;
            const bool ok = idx < total && mm < g.M;
            slot[r] = idx < total ? (ww * 3 + comp) * Ny + n : -1;
            // element 0 stands in for an element past the block (then zeroed),
            // so that every load is issued unconditionally
            const long off = ok ? ((long)comp * Ny + n) * g.M + mm : 0;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                v[r][j] = in[j][off];
                if (!ok) v[r][j] = dvec2{0.0, 0.0};
            }
        }
#pragma unroll
        for (int r = 0; r < UNR; ++r) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                acc.x += coef[j] * v[r][j].x;
                acc.y += coef[j] * v[r][j].y;
            }
            if (slot[r] >= 0) lds[slot[r]] = acc;
        }
    }
}

// waves per SIMD a tau-solve variant is compiled for: two 16-lane coefficient
// blocks fit in the 168 registers of three waves
constexpr int tau_waves(int K, int W) { return (W == 16 && K == 2) ? 3 : 1; }
constexpr int NXCD = 8;

// Block order of a launch over (mode block, member) pairs: the dispatcher
// hands consecutive workgroups to the 8 XCDs in turn, so the linear block
// index is renumbered to put a contiguous range of pairs on each XCD, members
// fastest. The members of a launch share their tau factors, which one XCD
// then reads from its own L2 for every member after the first.
// The launch's block (x, y, z) renumbered so that each XCD runs a contiguous
// range of the linear block order (x fastest): the dispatcher deals
// consecutive workgroups to the NXCD XCDs in turn, so without it neighbouring
// blocks, which share the cache lines at their edges, fetch those lines into
// different L2s. A permutation of the blocks: the work of each is unchanged.
struct BlockId {
    int x, y, z;
};
__device__ __forceinline__ BlockId xcd_grid() {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned T = gx * gy * gridDim.z;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = T / NXCD, r = T % NXCD, x = L % NXCD;
    const unsigned p = x * q + (x < r ? x : r) + L / NXCD;
    const unsigned t = p / gx;
    return BlockId{(int)(p - t * gx), (int)(t % gy), (int)(t / gy)};
}

__device__ __forceinline__ void xcd_block(int nmem, int& mode_block, int& member) {
    const int T = gridDim.x;  // mode blocks x members
    const int L = blockIdx.x;
    const int q = T / NXCD, r = T % NXCD, x = L % NXCD;
    const int p = x * q + (x < r ? x : r) + L / NXCD;
    mode_block = p / nmem;
    member = p - mode_block * nmem;
}

// Block of MPB consecutive modes of one member. Each wave solves its modes in
// registers from the right-hand sides staged in LDS, and u, v, w, q leave
// through LDS the same way.
template <int K, int W, int MPB>
__global__ void __launch_bounds__(W * MPB) __attribute__((amdgpu_waves_per_eu(tau_waves(K, W))))
k_tau_solve_wave(Geom g, const double* __restrict__ hb, int nmem, const TauMember* __restrict__ tab) {
    extern __shared__ double2 lds[];  // [MPB][3][Ny] right-hand sides, then the B rows [3][Ny]
    int mblk, mem;
    xcd_block(nmem, mblk, mem);
    const TauMember& me = tab[mem];
    const TauSet s = me.set;
    const double nu = me.nu;
    const int taucorr = me.taucorr;
    const SolveConst zc = me.sc;
    __attribute__((address_space(1))) dvec2* __restrict__ u = gmem_w(reinterpret_cast<dvec2*>(me.u));
    __attribute__((address_space(1))) dvec2* __restrict__ q = gmem_w(reinterpret_cast<dvec2*>(me.q));
    const int Ny = g.Ny;
    const int nthr = W * MPB;
    const int m0 = mblk * MPB;
    const int w = threadIdx.x / W;  // mode slot of this lane's segment
    const int m = m0 + w;
    const bool valid = m < g.M;
    double* hbl = reinterpret_cast<double*>(lds + 3 * Ny * MPB);
    for (int i = threadIdx.x; i < 3 * Ny; i += nthr) hbl[i] = hb[i];
    switch (me.nterms) {
        case 1: tau_stage_rhs<1, MPB>(me, g, m0, nthr, lds); break;
        case 2: tau_stage_rhs<2, MPB>(me, g, m0, nthr, lds); break;
        case 3: tau_stage_rhs<3, MPB>(me, g, m0, nthr, lds); break;
        case 4: tau_stage_rhs<4, MPB>(me, g, m0, nthr, lds); break;
        case 5: tau_stage_rhs<5, MPB>(me, g, m0, nthr, lds); break;
        case 6: tau_stage_rhs<6, MPB>(me, g, m0, nthr, lds); break;
        case 7: tau_stage_rhs<7, MPB>(me, g, m0, nthr, lds); break;
        default: tau_stage_rhs<MAX_TERMS, MPB>(me, g, m0, nthr, lds); break;
    }
    __syncthreads();
    WCtx c;
    c.N = Ny;
    c.Me = (Ny + 1) / 2;
    c.Mo = (Ny - 1) / 2;
    c.Nb = Ny - 1;
    c.Nc = Ny - 1;
    c.lane = threadIdx.x % W;
    c.L = g.b - g.a;
    c.hb = hbl;
    const int ikx = m % g.nkx, ikz = m / g.nkx;
    const int kx = kx_of(g, ikx), kz = ikz;
    const bool zero = (kx == 0 && kz == 0);
    const double pi2 = PI * PI;
    const double kk = (kx / g.Lx) * (kx / g.Lx) + (kz / g.Lz) * (kz / g.Lz);
    const double kappa2 = 4 * pi2 * kk;
    const double lambda = s.lambda_t + (4.0 * pi2 * nu) * kk;
    const double kx2pi = 2 * PI * kx / g.Lx, kz2pi = 2 * PI * kz / g.Lz;
    const gcdouble cm = gmem(s.coef) + (long)(valid ? m : 0) * TAU_NA * Ny;
    const gcdouble sc = gmem(s.scal) + (long)(valid ? m : 0) * 6;
    // LDS rows of this mode: Rx, Ry, Rz as staged. Each is read where it is
    // needed rather than held in registers across the solves (the registers
    // bound the waves per SIMD); the rows are then reused for the outputs.
    double2* const Sx = lds + (w * 3 + 0) * Ny;
    double2* const Sy = lds + (w * 3 + 1) * Ny;
    double2* const Sz = lds + (w * 3 + 2) * Ny;
    // Rx or Rz of this mode; kx = kz = 0 (nse.cpp:505-527) adds nu U'' (nu W'')
    // and, with a fixed pressure gradient, the mean gradient
    auto load_R = [&](const double2* S, const gcdouble base, double dPref, WV<K>& re, WV<K>& im) {
        wload_c<K, W>(S, c, re, im);
        if (zero) {
            const WV<K> Byy = wload<K, W>(base, c);
            for (int k = 0; k < K; ++k) {
                re.e[k] += Byy.e[k];
                re.o[k] += Byy.o[k];
            }
            if (!zc.bulk && c.lane == 0) re.e[0] -= dPref;
        }
    };
    const gcdouble nuUyy = gmem(zc.nuUbaseyy), nuWyy = gmem(zc.nuWbaseyy);
    WV<K> Pr, Pi, vr, vi;
    if (valid) {
        if (zero) {  // solve_P_and_v for kx = kz = 0: P from the pressure Helmholtz, v = 0
            WV<K> Ryr, Ryi;
            wload_c<K, W>(Sy, c, Ryr, Ryi);
            Pr = whelm<K, W>(wdiff<K, W>(Ryr, c), cm, cm + Ny, cm + 2L * Ny, kappa2, 0.0, 0.0, c);
            vr = wzero<K, W>();
            vi = wzero<K, W>();
            Pi = wzero<K, W>();  // real parts only (nse.cpp:555-563)
        } else {  // real and imaginary parts together
            WV<K> rr[2], P2[2], v2[2];
            {
                WV<K> Ry2[2];
                wload_c<K, W>(Sy, c, Ry2[0], Ry2[1]);
                wdiff_n<K, W, 2>(Ry2, c, rr);
            }
            {
                WV<K> Rxr, Rxi, Rzr, Rzi;
                wload_c<K, W>(Sx, c, Rxr, Rxi);
                wload_c<K, W>(Sz, c, Rzr, Rzi);
                for (int k = 0; k < K; ++k) {
                    rr[0].e[k] -= kx2pi * Rxi.e[k] + kz2pi * Rzi.e[k];
                    rr[0].o[k] -= kx2pi * Rxi.o[k] + kz2pi * Rzi.o[k];
                    rr[1].e[k] += kx2pi * Rxr.e[k] + kz2pi * Rzr.e[k];
                    rr[1].o[k] += kx2pi * Rxr.o[k] + kz2pi * Rzr.o[k];
                }
            }
            tau_P_and_v_wave2<K, W>(c, cm, sc, kappa2, lambda, nu, taucorr != 0, rr, Sy, P2, v2);
            Pr = P2[0];
            Pi = P2[1];
            vr = v2[0];
            vi = v2[1];
        }
        wstore_c<K, W>(Sy, c, vr, vi);  // Ry's row is free: v goes out through it
    }
    // rows [slot][n] of the block's modes -> global component array dst (stride M)
    auto drain = [&](int slot, __attribute__((address_space(1))) dvec2* dst) {
        for (int idx = threadIdx.x; idx < Ny * MPB; idx += nthr) {
            const int ww = idx % MPB, n = idx / MPB;
            const int mm = m0 + ww;
            if (mm < g.M) {
                const double2 v = lds[(ww * 3 + slot) * Ny + n];
                dst[(long)n * g.M + mm] = dvec2{v.x, v.y};
            }
        }
    };
    __syncthreads();
    drain(1, u + (long)Ny * g.M);  // v
    __syncthreads();
    if (valid) wstore_c<K, W>(Sy, c, Pr, Pi);
    __syncthreads();
    drain(1, q);  // q = P
    if (valid) {
        const gcdouble vb = cm + 3L * Ny;
        const gcdouble vu = cm + 4L * Ny;
        const gcdouble vd = cm + 5L * Ny;
        // u (row Sx) and w (row Sz) from the velocity Helmholtz operator. Called
        // with explicit register arrays: selecting between them at run time
        // would put them in scratch memory.
        auto solve_uw = [&](double k2, const WV<K>& Rr, const WV<K>& Ri, double umean, double* mu_out,
                            double2* row) {
            WV<K> fr, fi;
            for (int k = 0; k < K; ++k) {
                fr.e[k] = -(k2 * Pi.e[k]) - Rr.e[k];
                fr.o[k] = -(k2 * Pi.o[k]) - Rr.o[k];
                fi.e[k] = k2 * Pr.e[k] - Ri.e[k];
                fi.o[k] = k2 * Pr.o[k] - Ri.o[k];
            }
            WV<K> xr, xi;
            if (zc.bulk) {  // bulk-velocity constraint (tausolver.cpp:404-450)
                const double mu = whelm_mean<K, W>(fr, vb, vu, vd, lambda, nu, umean, c, xr);
                if (c.lane == 0) *mu_out = mu;
            } else {
                xr = whelm<K, W>(fr, vb, vu, vd, lambda, 0.0, 0.0, c);
            }
            xi = wzero<K, W>();
            wstore_c<K, W>(row, c, xr, xi);
        };
        if (zero) {  // kx = kz = 0: real parts only
            {
                WV<K> Rxr, Rxi;
                load_R(Sx, nuUyy, zc.dPdxRef, Rxr, Rxi);
                solve_uw(kx2pi, Rxr, Rxi, zc.umean, zc.dPdx_out, Sx);
            }
            WV<K> Rzr, Rzi;
            load_R(Sz, nuWyy, zc.dPdzRef, Rzr, Rzi);
            solve_uw(kz2pi, Rzr, Rzi, zc.wmean, zc.dPdz_out, Sz);
        } else {  // u and w, real and imaginary parts: four right-hand sides of one operator
            WV<K> f4[4], x4[4];
            {
                WV<K> Rxr, Rxi, Rzr, Rzi;
                wload_c<K, W>(Sx, c, Rxr, Rxi);
                wload_c<K, W>(Sz, c, Rzr, Rzi);
                for (int k = 0; k < K; ++k) {
                    f4[0].e[k] = -(kx2pi * Pi.e[k]) - Rxr.e[k];
                    f4[0].o[k] = -(kx2pi * Pi.o[k]) - Rxr.o[k];
                    f4[1].e[k] = kx2pi * Pr.e[k] - Rxi.e[k];
                    f4[1].o[k] = kx2pi * Pr.o[k] - Rxi.o[k];
                    f4[2].e[k] = -(kz2pi * Pi.e[k]) - Rzr.e[k];
                    f4[2].o[k] = -(kz2pi * Pi.o[k]) - Rzr.o[k];
                    f4[3].e[k] = kz2pi * Pr.e[k] - Rzi.e[k];
                    f4[3].o[k] = kz2pi * Pr.o[k] - Rzi.o[k];
                }
            }
            const double z4[4] = {0.0, 0.0, 0.0, 0.0};
            whelm_n<K, W, 4>(f4, vb, vu, vd, lambda, z4, z4, c, x4);
            wstore_c<K, W>(Sx, c, x4[0], x4[1]);
            wstore_c<K, W>(Sz, c, x4[2], x4[3]);
        }
    }
    __syncthreads();
    drain(0, u);                     // u
    drain(2, u + 2L * Ny * g.M);     // w
}

// PressureSolver::solve (poissonsolver.cpp:326-431) per mode, after the
// nonlinear term f of u + U: p'' - kappa^2 p = -div f with p = 0 at the walls,
// solved with the pressure factors of the tau data (the same Helmholtz operator,
// nu = 1), then for every mode but (0,0) the wall correction
// g = c exp(-mu (y-a)) + d exp(mu (y-b)), mu^2 = kappa^2, that makes
// dp/dy = nu v'' at y = a, b; g's coefficients are Ti times its values at the
// Chebyshev points.
template <int K, int W, int MPB>
__global__ void __launch_bounds__(W * MPB) k_pressure_wave(Geom g, YMats ym, PressArgs pa, const double* __restrict__ hb) {
    extern __shared__ double2 lds[];  // [MPB][4][Ny]: f0, f1, f2, v; then the B rows [3][Ny]
    const int Ny = g.Ny;
    double* hbl = reinterpret_cast<double*>(lds + 4 * Ny * MPB);
    for (int i = threadIdx.x; i < 3 * Ny; i += W * MPB) hbl[i] = hb[i];
    const long M = g.M;
    const int m0 = blockIdx.x * MPB, w = threadIdx.x / W;
    const long m = m0 + w;
    const bool valid = m < M;
    const double2* __restrict__ f = reinterpret_cast<const double2*>(pa.f);
    const double2* __restrict__ u = reinterpret_cast<const double2*>(pa.u);
    double2* __restrict__ q = reinterpret_cast<double2*>(pa.q);
    for (int idx = threadIdx.x; idx < 4 * Ny * MPB; idx += W * MPB) {
        const int ww = idx % MPB, rest = idx / MPB, n = rest % Ny, comp = rest / Ny;
        const long mm = m0 + ww;
        double2 v = make_double2(0.0, 0.0);
        if (mm < M) v = comp < 3 ? f[((long)comp * Ny + n) * M + mm] : u[((long)Ny + n) * M + mm];
        lds[(ww * 4 + comp) * Ny + n] = v;
    }
    __syncthreads();
    WCtx c;
    c.N = Ny;
    c.Me = (Ny + 1) / 2;
    c.Mo = (Ny - 1) / 2;
    c.Nb = Ny - 1;
    c.Nc = Ny - 1;
    c.lane = threadIdx.x % W;
    c.L = g.b - g.a;
    c.hb = hbl;
    if (valid) {
        const int ikx = (int)(m % g.nkx), ikz = (int)(m / g.nkx);
        const int kx = kx_of(g, ikx), kz = ikz;
        const bool zero = (kx == 0 && kz == 0);
        const double pi2 = PI * PI;
        const double kappa2 = 4 * pi2 * ((kx / g.Lx) * (kx / g.Lx) + (kz / g.Lz) * (kz / g.Lz));
        const double ax = 2 * PI * kx / g.Lx, az = 2 * PI * kz / g.Lz;
        const double* cm = pa.set.coef + m * TAU_NA * Ny;
        WV<K> Fxr, Fxi, Fyr, Fyi, Fzr, Fzi;
        wload_c<K, W>(lds + (w * 4 + 0) * Ny, c, Fxr, Fxi);
        wload_c<K, W>(lds + (w * 4 + 1) * Ny, c, Fyr, Fyi);
        wload_c<K, W>(lds + (w * 4 + 2) * Ny, c, Fzr, Fzi);
        const WV<K> Dr = wdiff<K, W>(Fyr, c), Di = wdiff<K, W>(Fyi, c);
        WV<K> Rr, Ri;  // -div f (diffops.cpp:2488-2516)
        for (int k = 0; k < K; ++k) {
            Rr.e[k] = -((-ax * Fxi.e[k] - az * Fzi.e[k]) + Dr.e[k]);
            Rr.o[k] = -((-ax * Fxi.o[k] - az * Fzi.o[k]) + Dr.o[k]);
            Ri.e[k] = -((ax * Fxr.e[k] + az * Fzr.e[k]) + Di.e[k]);
            Ri.o[k] = -((ax * Fxr.o[k] + az * Fzr.o[k]) + Di.o[k]);
        }
        WV<K> Pr = whelm<K, W>(Rr, cm, cm + Ny, cm + 2L * Ny, kappa2, 0.0, 0.0, c);
        WV<K> Pi = whelm<K, W>(Ri, cm, cm + Ny, cm + 2L * Ny, kappa2, 0.0, 0.0, c);
        if (!zero) {  // dp/dy = nu v'' at the walls (poissonsolver.cpp:356-427)
            WV<K> Vr, Vi;
            wload_c<K, W>(lds + (w * 4 + 3) * Ny, c, Vr, Vi);
            double pbr, par_, pbi, pai, vbr, var_, vbi, vai;
            wends<K, W>(wdiff<K, W>(Pr, c), c, pbr, par_);
            wends<K, W>(wdiff<K, W>(Pi, c), c, pbi, pai);
            wends<K, W>(wdiff<K, W>(wdiff<K, W>(Vr, c), c), c, vbr, var_);
            wends<K, W>(wdiff<K, W>(wdiff<K, W>(Vi, c), c), c, vbi, vai);
            const double alr = pa.nu * var_ - par_, ali = pa.nu * vai - pai;
            const double ber = pa.nu * vbr - pbr, bei = pa.nu * vbi - pbi;
            const double mu = sqrt(kappa2), H = g.b - g.a, e = exp(-mu * H);
            const double delta = mu * (1 - exp(-2 * mu * H));
            const double ccr = (-alr + ber * e) / delta, cci = (-ali + bei * e) / delta;
            const double ddr = (ber - alr * e) / delta, ddi = (bei - ali * e) / delta;
            const double piN = PI / (Ny - 1), rad = (g.b - g.a) / 2, cen = (g.b + g.a) / 2;  // chebypoints
            for (int par = 0; par < 2; ++par)
                for (int k = 0; k < K; ++k) {
                    const int n = 2 * (K * c.lane + k) + par;
                    if (n >= Ny) continue;
                    double sr = 0.0, si = 0.0;
                    for (int j = 0; j < Ny; ++j) {
                        const double y = cen + rad * cos(j * piN);
                        const double e1 = exp(-mu * (y - g.a)), e2 = exp(mu * (y - g.b));
                        const double t = ym.Ti[(long)n * Ny + j];
                        sr += t * (ccr * e1 + ddr * e2);
                        si += t * (cci * e1 + ddi * e2);
                    }
                    if (par == 0) {
                        Pr.e[k] += sr;
                        Pi.e[k] += si;
                    } else {
                        Pr.o[k] += sr;
                        Pi.o[k] += si;
                    }
                }
        }
        wstore_c<K, W>(lds + (w * 4 + 0) * Ny, c, Pr, Pi);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < Ny * MPB; idx += W * MPB) {
        const int ww = idx % MPB, n = idx / MPB;
        const long mm = m0 + ww;
        if (mm < M) q[(long)n * M + mm] = lds[(ww * 4 + 0) * Ny + n];
    }
}

// Raises a kernel's dynamic-LDS limit once per (kernel, size), so that launches
// recorded into a graph make no attribute calls.
template <class K>
static void allow_lds(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    static std::map<const void*, size_t> done;
    size_t& d = done[(const void*)kernel];
    if (d >= bytes) return;
    const hipError_t e =
        hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess)  // the launch that follows would fail without saying why
        throw std::runtime_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize, " + std::to_string(bytes) +
                                 " B): " + hipGetErrorString(e));
    d = bytes;
}

void launch_pressure(const Geom& g, const YMats& ym, const double* hb, const PressArgs& pa, hipStream_t st) {
    const int Me = (g.Ny + 1) / 2;
    auto go = [&](auto kern, int w) {
        const int mpb = 256 / w;
        const dim3 grid((unsigned)((g.M + mpb - 1) / mpb));
        const size_t lds = (size_t)4 * g.Ny * mpb * sizeof(double2) + (size_t)3 * g.Ny * sizeof(double);
        allow_lds(kern, lds);
        kern<<<grid, 256, lds, st>>>(g, ym, pa, hb);
    };
    if (Me <= 16) go(k_pressure_wave<1, 16, 16>, 16);
    else if (Me <= 32) go(k_pressure_wave<1, 32, 8>, 32);
    else if (Me <= 64) go(k_pressure_wave<1, 64, 4>, 64);
    else if (Me <= 128) go(k_pressure_wave<2, 64, 4>, 64);
    else go(k_pressure_wave<3, 64, 4>, 64);
}

void launch_tau_solve(const Geom& g, const double* hb, int n, const TauMember* tab, hipStream_t st) {
    const int Me = (g.Ny + 1) / 2;  // even coefficients: at most K per lane of a W-lane segment
    // 256-thread blocks: several per CU, so that one block's HBM staging overlaps
    // another's scans
    auto go = [&](auto kern, int w) {
        const int mpb = 256 / w;
        const long nblk = (long)((g.M + mpb - 1) / mpb) * n;
        if (nblk > 0x7fffffffL) throw std::runtime_error("tau solve: too many blocks in one launch");
        const size_t lds = (size_t)3 * g.Ny * mpb * sizeof(double2) + (size_t)3 * g.Ny * sizeof(double);
        allow_lds(kern, lds);
        kern<<<dim3((unsigned)nblk), 256, lds, st>>>(g, hb, n, tab);
    };
    // Up to 48 even coefficients: 16-lane segments (four modes per wave, every
    // shuffle a DPP row move) with up to three coefficients per lane. Above:
    // one or two modes per wave, shuffles through LDS permutes.
    if (Me <= 16) go(k_tau_solve_wave<1, 16, 16>, 16);
    else if (Me <= 32) go(k_tau_solve_wave<2, 16, 16>, 16);
    else if (Me <= 48) go(k_tau_solve_wave<3, 16, 16>, 16);
    else if (Me <= 64) go(k_tau_solve_wave<1, 64, 4>, 64);
    else if (Me <= 96) go(k_tau_solve_wave<3, 32, 8>, 32);
    else if (Me <= 128) go(k_tau_solve_wave<2, 64, 4>, 64);
    else go(k_tau_solve_wave<3, 64, 4>, 64);
}

// ===========================================================================
// Elementwise combination (dnsalgo.cpp:218-224, 668-673)
// ===========================================================================
struct CombineArgs {
    double coef[8];
    const double* in[8];
};

__global__ void k_combine(long n, int nterms, CombineArgs a, double* out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < nterms; ++j) acc += a.coef[j] * a.in[j][i];
    out[i] = acc;
}

void launch_combine(long n, int nterms, const double* coef, const double* const* in, double* out, hipStream_t st) {
    CombineArgs a;
    for (int j = 0; j < 8; ++j) {
        a.coef[j] = j < nterms ? coef[j] : 0.0;
        a.in[j] = j < nterms ? in[j] : nullptr;
    }
    const int blk = 256;
    k_combine<<<(int)((n + blk - 1) / blk), blk, 0, st>>>(n, nterms, a, out);
}

// ===========================================================================
// Finiteness of member states (the isfinite(L2Norm(u)) checks of
// cfdsi.cpp:707-710, 769-774): flags[m] = 1 if field m holds a NaN or an
// infinity. Blocks stride over the field; a wave with a non-finite value
// stores the flag (every such store writes the same 1).
// ===========================================================================
__global__ void __launch_bounds__(256) k_nonfinite(const double* const* __restrict__ fields, long n,
                                                   double* __restrict__ flags) {
    const double* __restrict__ f = fields[blockIdx.y];
    bool bad = false;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        bad = bad || !isfinite(f[i]);
    if (__any(bad) && bad) flags[blockIdx.y] = 1.0;
}

void launch_nonfinite(int nfields, const double* const* fields, long n, double* flags, hipStream_t st) {
    if (nfields <= 0) return;
    const int blk = 256;
    const long want = (n + blk * 8L - 1) / (blk * 8L);  // about 8 values per thread
    const unsigned gx = (unsigned)std::max(1L, std::min(want, 1024L));
    k_nonfinite<<<dim3(gx, (unsigned)nfields), blk, 0, st>>>(fields, n, flags);
}

// ===========================================================================
// NSE::linear on live modes (nse.cpp:393-477), one thread per mode
// ===========================================================================
// NSE::linear (nse.cpp:393-477) with one W-lane segment of a wavefront per
// mode, MPB modes per block, member blockIdx.y: for each velocity component
// L = nu (u'' - kappa^2 u) - grad q, plus on kx = kz = 0 nu U'', nu W'' and
// the mean pressure gradient (the bulk-velocity branch takes dPdx from
// nu (du/dy(b) - du/dy(a)) / Ly, nse.cpp:437-472). Real and imaginary parts
// share each derivative scan.
template <int K, int W, int MPB>
__global__ void __launch_bounds__(W * MPB) k_linear_wave(Geom g, const LinearMember* __restrict__ tab) {
    const LinearMember& me = tab[blockIdx.y];
    const int m = blockIdx.x * MPB + threadIdx.x / W;
    if (m >= g.M) return;  // whole segments leave together
    const double nu = me.nu;
    const LinearConst lc = me.c;
    const gcdouble u = gmem(me.u), q = gmem(me.q);
    __attribute__((address_space(1))) double* out = gmem_w(me.out);
    const int Ny = g.Ny;
    WCtx c;
    c.N = Ny;
    c.Me = (Ny + 1) / 2;
    c.Mo = (Ny - 1) / 2;
    c.Nb = Ny - 1;
    c.Nc = Ny - 1;
    c.lane = threadIdx.x % W;
    c.L = g.b - g.a;
    c.hb = nullptr;
    const int ikx = m % g.nkx, kz = m / g.nkx, kx = kx_of(g, ikx);
    const bool zero = kx == 0 && kz == 0;
    const double kappa2 = 4 * PI * PI * ((kx / g.Lx) * (kx / g.Lx) + (kz / g.Lz) * (kz / g.Lz));
    const long rs = 2L * g.M;  // doubles between consecutive Chebyshev rows
    auto row0 = [&](int comp) { return 2L * ((long)comp * Ny * g.M + m); };
    WV<K> Pq[2] = {wload_s<K, W>(q + row0(0), rs, c), wload_s<K, W>(q + row0(0) + 1, rs, c)};
    double dPd[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        WV<K> s0[2] = {wload_s<K, W>(u + row0(comp), rs, c), wload_s<K, W>(u + row0(comp) + 1, rs, c)};
        for (int p = 0; p < 2; ++p)
            for (int k = 0; k < K; ++k) {
                s0[p].e[k] *= nu;  // nu u
                s0[p].o[k] *= nu;
            }
        WV<K> s1[2], s2[2];
        wdiff_n<K, W, 2>(s0, c, s1);
        if (zero && lc.bulk && comp != 1) {
            double eb, ea;
            wends<K, W>(s1[0], c, eb, ea);
            dPd[comp] = (eb - ea) / c.L;
        }
        wdiff_n<K, W, 2>(s1, c, s2);  // nu u''
        WV<K> o[2];
        for (int p = 0; p < 2; ++p)
            for (int k = 0; k < K; ++k) {
                o[p].e[k] = s2[p].e[k] - kappa2 * s0[p].e[k];
                o[p].o[k] = s2[p].o[k] - kappa2 * s0[p].o[k];
            }
        if (comp == 1) {  // - dq/dy
            WV<K> dq[2];
            wdiff_n<K, W, 2>(Pq, c, dq);
            for (int p = 0; p < 2; ++p)
                for (int k = 0; k < K; ++k) {
                    o[p].e[k] = o[p].e[k] - dq[p].e[k];
                    o[p].o[k] = o[p].o[k] - dq[p].o[k];
                }
        } else {  // - (i k) q: (0, d) * (Pr, Pi) = (-(d Pi), d Pr)
            const double d = comp == 0 ? 2 * PI * kx / g.Lx : 2 * PI * kz / g.Lz;
            for (int k = 0; k < K; ++k) {
                o[0].e[k] = o[0].e[k] - (0.0 * Pq[0].e[k] - d * Pq[1].e[k]);
                o[0].o[k] = o[0].o[k] - (0.0 * Pq[0].o[k] - d * Pq[1].o[k]);
                o[1].e[k] = o[1].e[k] - (0.0 * Pq[1].e[k] + d * Pq[0].e[k]);
                o[1].o[k] = o[1].o[k] - (0.0 * Pq[1].o[k] + d * Pq[0].o[k]);
            }
        }
        if (zero && comp != 1) {
            const WV<K> base = wload<K, W>(gmem(comp == 0 ? lc.nuUbaseyy : lc.nuWbaseyy), c);
            for (int k = 0; k < K; ++k) {
                o[0].e[k] += base.e[k];
                o[0].o[k] += base.o[k];
            }
            if (c.lane == 0) {
                if (!lc.bulk) o[0].e[0] -= comp == 0 ? lc.dPdxRef : lc.dPdzRef;
                else o[0].e[0] -= dPd[comp] + (comp == 0 ? lc.nuUbase_slope : lc.nuWbase_slope);
            }
        }
        wstore_s<K, W>(out + row0(comp), rs, o[0], c);
        wstore_s<K, W>(out + row0(comp) + 1, rs, o[1], c);
    }
}

void launch_linear(const Geom& g, int n, const LinearMember* tab, hipStream_t st) {
    const int Me = (g.Ny + 1) / 2;
    auto go = [&](auto kern, int w) {
        const int mpb = 256 / w;
        kern<<<dim3((g.M + mpb - 1) / mpb, n), 256, 0, st>>>(g, tab);
    };
    if (Me <= 16) go(k_linear_wave<1, 16, 16>, 16);
    else if (Me <= 32) go(k_linear_wave<2, 16, 16>, 16);
    else if (Me <= 48) go(k_linear_wave<3, 16, 16>, 16);
    else if (Me <= 64) go(k_linear_wave<1, 64, 4>, 64);
    else if (Me <= 96) go(k_linear_wave<3, 32, 8>, 32);
    else if (Me <= 128) go(k_linear_wave<2, 64, 4>, 64);
    else if (Me <= 192) go(k_linear_wave<3, 64, 4>, 64);
    else throw std::runtime_error("NSE::linear: Ny above 383 is not supported");
}

// ===========================================================================
// Nonlinear term
// ===========================================================================
// inverse x-FFT: yp [F][Ny][nkz][nkx] -> xp [Fo][Ny][Nx][nkz]; output f >= nsrc is
// d/dx of source field f - nsrc. XL kz-lines per block.
template <int NX>
__global__ void __launch_bounds__(256) k_xinv(Geom g, FFTPlan px, int XL, int nsrc, const double2* __restrict__ yp0,
                                              long sy, double2* __restrict__ xp0, long sx) {
    extern __shared__ double2 lds[];
    const double2* __restrict__ yp = yp0 + blockIdx.z * sy;
    double2* __restrict__ xp = xp0 + blockIdx.z * sx;
    const int Nx = g.Nx, nkz = g.nkz, nkx = g.nkx, Ny = g.Ny;
    const int nchunk = (nkz + XL - 1) / XL;
    const int y = blockIdx.x / nchunk, ikz0 = (blockIdx.x % nchunk) * XL;
    const int fo = blockIdx.y;
    const bool deriv = fo >= nsrc;
    const int fs = deriv ? fo - nsrc : fo;
    const int nl = min(XL, nkz - ikz0);
    const int P = lpitch(Nx);
    double2* tw = lds;
    double2* a = lds + Nx;
    double2* b = a + XL * P;
    stage_tw(tw, px);
    const double dk = 2 * PI / g.Lx;
    for (int w = threadIdx.x; w < nl * Nx; w += blockDim.x) {
        const int l = w / Nx, j = w - l * Nx;
        int ikx = -1;
        if (j <= g.kxd) ikx = j;
        else if (j >= Nx - g.kxd) ikx = j - Nx + nkx;
        double2 v = make_double2(0.0, 0.0);
        if (ikx >= 0) {
            v = yp[(((long)fs * Ny + y) * nkz + ikz0 + l) * nkx + ikx];
            if (deriv) v = imul(dk * kx_of(g, ikx), v);
        }
        a[l * P + pidx(j)] = v;
    }
    __syncthreads();
    double2* r = fft_any<NX, +1>(a, b, px, tw, nl);
    double2* out = xp + (((long)fo * Ny + y) * Nx) * nkz + ikz0;
    for (int w = threadIdx.x; w < nl * Nx; w += blockDim.x) {
        const int x = w / nl, l = w - x * nl;
        out[(long)x * nkz + l] = r[l * P + pidx(x)];
    }
}

// forward x-FFT: zo [G][Ny][Nx][nkz] -> fo [G][Ny][nkz][nkx], scaled by 1/(Nx Nz)
template <int NX>
__global__ void __launch_bounds__(256) k_xfwd(Geom g, FFTPlan px, int XL, const double2* __restrict__ zo0, long sx,
                                              double2* __restrict__ fo0, long sy) {
    extern __shared__ double2 lds[];
    const double2* __restrict__ zo = zo0 + blockIdx.z * sx;
    double2* __restrict__ fo = fo0 + blockIdx.z * sy;
    const int Nx = g.Nx, nkz = g.nkz, nkx = g.nkx, Ny = g.Ny;
    const int nchunk = (nkz + XL - 1) / XL;
    const int y = blockIdx.x / nchunk, ikz0 = (blockIdx.x % nchunk) * XL;
    const int f = blockIdx.y;
    const int nl = min(XL, nkz - ikz0);
    const int P = lpitch(Nx);
    double2* tw = lds;
    double2* a = lds + Nx;
    double2* b = a + XL * P;
    stage_tw(tw, px);
    const double2* in = zo + (((long)f * Ny + y) * Nx) * nkz + ikz0;
    for (int w = threadIdx.x; w < nl * Nx; w += blockDim.x) {
        const int x = w / nl, l = w - x * nl;
        a[l * P + pidx(x)] = in[(long)x * nkz + l];
    }
    __syncthreads();
    double2* r = fft_any<NX, -1>(a, b, px, tw, nl);
    const double scale = 1.0 / ((double)g.Nx * g.Nz);
    for (int w = threadIdx.x; w < nl * nkx; w += blockDim.x) {
        const int l = w / nkx, ikx = w - l * nkx;
        const int j = ikx <= g.kxd ? ikx : ikx - nkx + Nx;
        const double2 v = r[l * P + pidx(j)];
        fo[(((long)f * Ny + y) * nkz + ikz0 + l) * nkx + ikx] = make_double2(v.x * scale, v.y * scale);
    }
}

// ---- z pass --------------------------------------------------------------
struct ZCtx {
    Geom g;
    FFTPlan pz;
    int y, x0, nl;
    double2 *wa, *wb, *tw;
    double* phys;  // [NP][XB][Nz]
    int XB;
};

__device__ __forceinline__ double* physf(const ZCtx& c, int f) { return c.phys + (long)f * c.XB * c.g.Nz; }

// c2r of two real fields at once: field fa -> phys[da], fb -> phys[db]
// (fb < 0: none). dza/dzb apply d/dz. Imag parts of kz = 0 are ignored, as
// FFTW's c2r does. The unused middle bins are zeroed in the same sweep.
template <int NZ>
__device__ void zc2r_pair(const ZCtx& c, const double2* __restrict__ xp, int fa, bool dza, int da, int fb, bool dzb,
                          int db) {
    const Geom& g = c.g;
    const int Nz = g.Nz, nkz = g.nkz;
    const int nmid = Nz - 2 * nkz + 1;  // bins nkz .. Nz-nkz
    const int P = lpitch(Nz);
    for (int w = threadIdx.x; w < c.nl * nmid; w += blockDim.x) {
        const int l = w / nmid, j = w - l * nmid;
        c.wa[l * P + pidx(nkz + j)] = make_double2(0.0, 0.0);
    }
    const long fsz = (long)g.Ny * g.Nx * nkz;
    const double dk = 2 * PI / g.Lz;
    for (int w = threadIdx.x; w < c.nl * nkz; w += blockDim.x) {
        const int l = w / nkz, k = w - l * nkz;
        const long base = ((long)c.y * g.Nx + c.x0 + l) * nkz + k;
        double2 A = xp[(long)fa * fsz + base];
        if (dza) A = imul(dk * k, A);
        double2 B = make_double2(0.0, 0.0);
        if (fb >= 0) {
            B = xp[(long)fb * fsz + base];
            if (dzb) B = imul(dk * k, B);
        }
        double2* z = c.wa + l * P;
        if (k == 0) {
            z[0] = make_double2(A.x, B.x);
        } else {
            z[pidx(k)] = make_double2(A.x - B.y, A.y + B.x);       // A + iB
            z[pidx(Nz - k)] = make_double2(A.x + B.y, B.x - A.y);  // conj(A) + i conj(B)
        }
    }
    __syncthreads();
    double2* r = fft_any<NZ, +1, true>(c.wa, c.wb, c.pz, c.tw, c.nl);
    double* pa = physf(c, da);
    double* pb = db >= 0 ? physf(c, db) : nullptr;
    for (int w = threadIdx.x; w < c.nl * Nz; w += blockDim.x) {
        const int l = w / Nz, zz = w - l * Nz;
        const double2 v = r[l * P + pidx(zz)];
        pa[w] = v.x;
        if (pb) pb[w] = v.y;
    }
    __syncthreads();
}

// r2c of phys[sa] (and phys[sb]) -> zo fields ga, gb (kz <= kzd kept).
// fold != 0 (with sb >= 0): field ga receives A + fold i kz B instead of A,
// kz = 2 pi k / Lz, A and B the spectra of phys[sa] and phys[sb]; gb < 0 then
// drops B.
template <int NZ>
__device__ void zr2c_pair(const ZCtx& c, double2* __restrict__ zo, int sa, int ga, int sb, int gb,
                          double fold = 0.0) {
    const Geom& g = c.g;
    const int Nz = g.Nz, nkz = g.nkz;
    const double* pa = physf(c, sa);
    const double* pb = sb >= 0 ? physf(c, sb) : nullptr;
    const int P = lpitch(Nz);
    for (int w = threadIdx.x; w < c.nl * Nz; w += blockDim.x) {
        const int l = w / Nz, zz = w - l * Nz;
        c.wa[l * P + pidx(zz)] = make_double2(pa[w], pb ? pb[w] : 0.0);
    }
    __syncthreads();
    double2* r = fft_any<NZ, -1, true>(c.wa, c.wb, c.pz, c.tw, c.nl);
    const long fsz = (long)g.Ny * g.Nx * nkz;
    for (int w = threadIdx.x; w < c.nl * nkz; w += blockDim.x) {
        const int l = w / nkz, k = w - l * nkz;
        const double2 Zk = r[l * P + pidx(k)];
        const double2 Zn = r[l * P + pidx((Nz - k) % Nz)];
        const long base = ((long)c.y * g.Nx + c.x0 + l) * nkz + k;
        const double2 A = make_double2(0.5 * (Zk.x + Zn.x), 0.5 * (Zk.y - Zn.y));
        const double2 B = make_double2(0.5 * (Zk.y + Zn.y), -0.5 * (Zk.x - Zn.x));
        if (fold != 0.0 && pb) {
            const double fk = fold * ((2 * PI / g.Lz) * k);
            zo[(long)ga * fsz + base] = make_double2(A.x - fk * B.y, A.y + fk * B.x);
        } else {
            zo[(long)ga * fsz + base] = A;
        }
        if (gb >= 0) zo[(long)gb * fsz + base] = B;
    }
    __syncthreads();
}

// physical-space products. xp fields: ROT u v w wx wy wz; CONV/SKEW u v w
// du/dy dv/dy dw/dy du/dx dv/dx dw/dx; DIV u v w. zo outputs: g0..g2 (the
// u.grad u / omega x u part and the rotation term) then uu00 uu01 uu02 uu11
// uu12 uu22 for DIV; for SKEW g_i + i kz u_i u_2 / 2 then uu00 uu01 uu02 uu11
// uu12 (8 fields).
template <int METHOD, int NZ>
__global__ void __launch_bounds__(256) k_zpass(Geom g, FFTPlan pz, int XB, const NLMember* __restrict__ tab,
                                               int want_g, const double2* __restrict__ xp0, double2* __restrict__ zo0,
                                               long sx) {
    extern __shared__ double2 lds[];
    const double rotation = tab[blockIdx.y].rotation;
    const double2* __restrict__ xp = xp0 + blockIdx.y * sx;
    double2* __restrict__ zo = zo0 + blockIdx.y * sx;
    const int nxc = (g.Nx + XB - 1) / XB;
    ZCtx c;
    c.g = g;
    c.pz = pz;
    c.y = blockIdx.x / nxc;
    c.x0 = (blockIdx.x % nxc) * XB;
    c.nl = min(XB, g.Nx - c.x0);
    c.XB = XB;
    c.tw = lds;
    c.wa = lds + g.Nz;
    c.wb = c.wa + XB * lpitch(g.Nz);
    c.phys = reinterpret_cast<double*>(c.wb + XB * lpitch(g.Nz));
    stage_tw(c.tw, pz);
    const int Nz = g.Nz;
    const int npts = c.nl * Nz;
    if (METHOD == NL_ROT) {
        zc2r_pair<NZ>(c, xp, 0, false, 0, 1, false, 1);
        zc2r_pair<NZ>(c, xp, 2, false, 2, 3, false, 3);
        zc2r_pair<NZ>(c, xp, 4, false, 4, 5, false, 5);
        double *u0 = physf(c, 0), *u1 = physf(c, 1), *u2 = physf(c, 2);
        double *w0 = physf(c, 3), *w1 = physf(c, 4), *w2 = physf(c, 5);
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {  // cross(vort, u) (diffops.cpp:2598-2605)
            const double a0 = u0[w], a1 = u1[w], a2 = u2[w], b0 = w0[w], b1 = w1[w], b2 = w2[w];
            double f0 = b1 * a2 - b2 * a1;
            double f1 = b2 * a0 - b0 * a2;
            const double f2 = b0 * a1 - b1 * a0;
            if (rotation != 0.0) {  // nse.cpp:61-80
                f0 -= rotation * a1;
                f1 += rotation * a0;
            }
            w0[w] = f0;
            w1[w] = f1;
            w2[w] = f2;
        }
        __syncthreads();
        zr2c_pair<NZ>(c, zo, 3, 0, 4, 1);
        zr2c_pair<NZ>(c, zo, 5, 2, -1, -1);
        return;
    }
    double *u0 = physf(c, 0), *u1 = physf(c, 1), *u2 = physf(c, 2);
    double *pa = physf(c, 3), *pb = physf(c, 4);
    if constexpr (METHOD == NL_CONV || METHOD == NL_SKEW) {
        // f_i = sum_j (s u_j) du_i/dx_j with the terms of each f_i added in the
        // order j = x, y, z (diffops.cpp:3192-3199, 3626-3633). Component by
        // component through phys 3, 4 (inputs) and 5, 6 (f); the inputs are
        // paired so that every two-for-one c2r is full.
        constexpr bool SK = (METHOD == NL_SKEW);
        const double s = SK ? 0.5 : 1.0;
        double *fa = physf(c, 5), *fb = physf(c, 6);
        auto su = [&](const double* uj, int w) { return SK ? s * uj[w] : uj[w]; };
        zc2r_pair<NZ>(c, xp, 0, false, 0, 1, false, 1);  // u, v
        zc2r_pair<NZ>(c, xp, 2, false, 2, 6, false, 3);  // w, du/dx
        for (int w = threadIdx.x; w < npts; w += blockDim.x) fa[w] = 0.0 + su(u0, w) * pa[w];
        zc2r_pair<NZ>(c, xp, 3, false, 3, 0, true, 4);   // du/dy, du/dz
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            fa[w] += su(u1, w) * pa[w];
            fa[w] += su(u2, w) * pb[w];
            if (rotation != 0.0) fa[w] -= rotation * u1[w];  // nse.cpp:61-80
            if (SK) pb[w] = u0[w] * u2[w];
        }
        __syncthreads();
        // skew-symmetric: f0 + i kz u0u2 / 2 and u0u2 (the layout of k_zwave)
        zr2c_pair<NZ>(c, zo, 5, 0, SK ? 4 : -1, SK ? 5 : -1, SK ? 0.5 : 0.0);
        zc2r_pair<NZ>(c, xp, 7, false, 3, 4, false, 4);  // dv/dx, dv/dy
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            fa[w] = 0.0 + su(u0, w) * pa[w];
            fa[w] += su(u1, w) * pb[w];
        }
        zc2r_pair<NZ>(c, xp, 1, true, 3, 8, false, 4);   // dv/dz, dw/dx
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            fa[w] += su(u2, w) * pa[w];
            if (rotation != 0.0) fa[w] += rotation * u0[w];
            fb[w] = 0.0 + su(u0, w) * pb[w];
        }
        zc2r_pair<NZ>(c, xp, 5, false, 3, 2, true, 4);   // dw/dy, dw/dz
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            fb[w] += su(u1, w) * pa[w];
            fb[w] += su(u2, w) * pb[w];
        }
        __syncthreads();
        if (SK) {
            // f1 + i kz u1u2 / 2 and u1u2, f2 + i kz u2u2 / 2, then the rest of
            // outer(u, u) (diffops.cpp:2336-2388): u0u0, u0u1, u1u1
            for (int w = threadIdx.x; w < npts; w += blockDim.x) {
                pa[w] = u1[w] * u2[w];
                pb[w] = u2[w] * u2[w];
            }
            __syncthreads();
            zr2c_pair<NZ>(c, zo, 5, 1, 3, 7, 0.5);
            zr2c_pair<NZ>(c, zo, 6, 2, 4, -1, 0.5);
            for (int w = threadIdx.x; w < npts; w += blockDim.x) {
                pa[w] = u0[w] * u0[w];
                pb[w] = u0[w] * u1[w];
            }
            __syncthreads();
            zr2c_pair<NZ>(c, zo, 3, 3, 4, 4);
            for (int w = threadIdx.x; w < npts; w += blockDim.x) pa[w] = u1[w] * u1[w];
            __syncthreads();
            zr2c_pair<NZ>(c, zo, 3, 6, -1, -1);
        } else {
            zr2c_pair<NZ>(c, zo, 5, 1, 6, 2);  // f1, f2
        }
    } else {
        zc2r_pair<NZ>(c, xp, 0, false, 0, 1, false, 1);
        zc2r_pair<NZ>(c, xp, 2, false, 2, -1, false, -1);
        if (want_g) {  // DIV with rotation: g = rotation terms only
            for (int w = threadIdx.x; w < npts; w += blockDim.x) {
                pa[w] = -(rotation * u1[w]);
                pb[w] = rotation * u0[w];
            }
            __syncthreads();
            zr2c_pair<NZ>(c, zo, 3, 0, 4, 1);
            for (int w = threadIdx.x; w < npts; w += blockDim.x) pa[w] = 0.0;
            __syncthreads();
            zr2c_pair<NZ>(c, zo, 3, 2, -1, -1);
        }
        // outer(u, u) (diffops.cpp:2336-2388), two entries per r2c
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            pa[w] = u0[w] * u0[w];
            pb[w] = u0[w] * u1[w];
        }
        __syncthreads();
        zr2c_pair<NZ>(c, zo, 3, 3, 4, 4);
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            pa[w] = u0[w] * u2[w];
            pb[w] = u1[w] * u1[w];
        }
        __syncthreads();
        zr2c_pair<NZ>(c, zo, 3, 5, 4, 6);
        for (int w = threadIdx.x; w < npts; w += blockDim.x) {
            pa[w] = u1[w] * u2[w];
            pb[w] = u2[w] * u2[w];
        }
        __syncthreads();
        zr2c_pair<NZ>(c, zo, 3, 7, 4, 8);
    }
}

// ---- wave-local z pass -------------------------------------------------------
// A z-line of length N = N1 N2 lives on G lanes of one wavefront (G the power
// of two >= N1, N2; 64/G lines per wavefront), in one of two layouts:
//   spectral: lane j < N1 holds Z[j + N1 r], r < N2;
//   physical: lane j < N2 holds x[j + N2 m], m < N1.
// The inverse transform is an N2-point DFT in registers, the twiddles
// w_N^(j k1), a transpose through the line's own LDS rows and an N1-point DFT
// in registers; the forward transform runs the same steps the other way. All
// fields of a line share the physical layout, so the products are lane-local,
// and a wavefront never waits for another one.
template <int NZ>
struct ZWShape {
    static constexpr int N1 = 0, N2 = 0;
};
template <> struct ZWShape<16> { static constexpr int N1 = 4, N2 = 4; };
template <> struct ZWShape<32> { static constexpr int N1 = 4, N2 = 8; };
template <> struct ZWShape<48> { static constexpr int N1 = 6, N2 = 8; };
template <> struct ZWShape<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct ZWShape<96> { static constexpr int N1 = 8, N2 = 12; };
template <> struct ZWShape<128> { static constexpr int N1 = 8, N2 = 16; };
template <> struct ZWShape<192> { static constexpr int N1 = 12, N2 = 16; };
template <> struct ZWShape<256> { static constexpr int N1 = 16, N2 = 16; };

constexpr int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

template <int NZ>
struct ZW {
    static constexpr int N1 = ZWShape<NZ>::N1, N2 = ZWShape<NZ>::N2, N = N1 * N2;
    static constexpr int NMAX = N1 > N2 ? N1 : N2;
    static constexpr int G = pow2_at_least(NMAX);  // lanes per line
    static constexpr int LPW = 64 / G;             // lines per wavefront
    static constexpr int LBUF = N + NMAX;          // LDS slots per line (transposes, padded rows)
};

constexpr bool zwave_size(int n) {
    return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 192 || n == 256;
}

// Orders the LDS accesses of one wavefront (its lanes read what other lanes of
// the same wavefront wrote): a compiler barrier between wavefront-scope fences.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- DFTs held in registers: sum_n v_n exp(SIGN 2 pi i n k / R) ----
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(double2 (&v)[R]);

constexpr int inv_mod(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 1;
}

// prime-factor (Good-Thomas) DFT of length A B with gcd(A, B) = 1: input
// n = (B a + A b) mod AB, output k with k = ka (mod A), k = kb (mod B); no twiddles
template <int A, int B, int SIGN>
__device__ __forceinline__ void dft_pfa(double2 (&v)[A * B]) {
    constexpr int R = A * B, BI = inv_mod(B, A), AI = inv_mod(A, B);
    double2 t[B][A];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        double2 s[A];
#pragma unroll
        for (int a = 0; a < A; ++a) s[a] = v[(B * a + A * b) % R];
        dft_reg<A, SIGN>(s);
#pragma unroll
        for (int a = 0; a < A; ++a) t[b][a] = s[a];
    }
#pragma unroll
    for (int ka = 0; ka < A; ++ka) {
        double2 s[B];
#pragma unroll
        for (int b = 0; b < B; ++b) s[b] = t[b][ka];
        dft_reg<B, SIGN>(s);
#pragma unroll
        for (int kb = 0; kb < B; ++kb) v[(B * BI * ka + A * AI * kb) % R] = s[kb];
    }
}

// exp(i m pi / 8), m <= 9, from cos(pi/8), sin(pi/8) and 1/sqrt(2)
__device__ __forceinline__ constexpr double2 w16(int m) {
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    return m == 1 ? double2{c1, s1} : m == 2 ? double2{h, h} : m == 3 ? double2{s1, c1} : m == 4 ? double2{0.0, 1.0}
         : m == 5 ? double2{-s1, c1} : m == 6 ? double2{-h, h} : m == 7 ? double2{-c1, s1} : m == 8 ? double2{-1.0, 0.0}
         : m == 9 ? double2{-c1, -s1} : double2{1.0, 0.0};
}

// length 16 as 4 x 4 (Cooley-Tukey, twiddles exp(SIGN 2 pi i n2 k1 / 16))
template <int SIGN>
__device__ __forceinline__ void dft_ct44(double2 (&v)[16]) {
    double2 t[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        double2 s[4];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) s[n1] = v[4 * n1 + n2];
        dft_small<4, SIGN>(s);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            if (n2 * k1 == 0) {
                t[n2][k1] = s[k1];
            } else {
                double2 w = w16(n2 * k1);
                if (SIGN < 0) w.y = -w.y;
                t[n2][k1] = cmul(s[k1], w);
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        double2 s[4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) s[n2] = t[n2][k1];
        dft_small<4, SIGN>(s);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = s[k2];
    }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(double2 (&v)[R]) {
    if constexpr (R == 2 || R == 3 || R == 4) {
        double2 t[4];
#pragma unroll
        for (int i = 0; i < R; ++i) t[i] = v[i];
        dft_small<R, SIGN>(t);
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = t[i];
    } else if constexpr (R == 8) {
        dft_small<8, SIGN>(v);
    } else if constexpr (R == 6) {
        dft_pfa<2, 3, SIGN>(v);
    } else if constexpr (R == 12) {
        dft_pfa<4, 3, SIGN>(v);
    } else if constexpr (R == 16) {
        dft_ct44<SIGN>(v);
    } else {
        static_assert(R == 0, "no register DFT of this length");
    }
}

// One line's view: its lane in the group, where its data sit, its LDS rows.
struct ZWCtx {
    int gi;            // lane within the line's group
    bool line_ok;      // the line exists (x < Nx)
    long base, fsz;    // line offset and field size in the [field][y][x][kz] arrays
    int nkz;
    double dk;         // 2 pi / Lz
    double2* T;        // the line's LDS rows
    const double2* tw; // exp(-2 pi i m / N), m < N, in LDS
    // input staging: the wavefront's lines are one contiguous run of the
    // [field][y][x][kz] arrays (wcount values from wbase); two fields of it sit
    // in LDS (slots 0 and 1 from S, slot stride sstride), this line's bins at
    // L * nkz within a slot
    long wbase;
    int wcount, lane, sline, sstride;
    double2* S;
};

// Starts the copy of field f of the wavefront's lines into staging slot
// `slot` (global_load_lds_dwordx4: lane i of a copy instruction moves 16 bytes
// to its base + 16 i; lanes past the run copy nothing). The data are in LDS
// once the wavefront's vector-memory count has drained; the compiler waits for
// it before the first LDS read of the staging array.
template <int NZ>
__device__ __forceinline__ void zw_stage(const ZWCtx& c, const double2* __restrict__ xp, int f, int slot) {
    const __attribute__((address_space(1))) double2* src = gmem(xp) + (long)f * c.fsz + c.wbase;
    double2* dst = c.S + slot * c.sstride;
    for (int t = 0; t < c.wcount; t += 64) {
        if (t + c.lane < c.wcount)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + t + c.lane),
                                             (__attribute__((address_space(3))) void*)(dst + t), 16, 0, 0);
    }
}

// DFT (sign SIGN) of a line held in layout A, v[r] = Z[j + N1 r] on lane
// j < N1, into layout B, u[m] = X[j + N2 m] on lane j < N2:
// X[k1 + N2 k2] = sum_j w_N1^(j k2) w_N^(j k1) sum_r Z[j + N1 r] w_N2^(r k1).
template <int N, int SIGN>
__device__ __forceinline__ void zw_a2b(double2 (&v)[ZW<N>::N2], double2 (&u)[ZW<N>::N1], int j, double2* T,
                                       const double2* tw) {
    constexpr int N1 = ZW<N>::N1, N2 = ZW<N>::N2;
    dft_reg<N2, SIGN>(v);
    const int jt = j < N1 ? j : N1 - 1;
#pragma unroll
    for (int k1 = 1; k1 < N2; ++k1) {
        double2 t = tw[jt * k1];
        if (SIGN > 0) t.y = -t.y;
        v[k1] = cmul(v[k1], t);
    }
    wave_lds_sync();
    if (j < N1) {
#pragma unroll
        for (int k1 = 0; k1 < N2; ++k1) T[k1 * (N1 + 1) + j] = v[k1];
    }
    wave_lds_sync();
    const int row = j < N2 ? j : N2 - 1;
#pragma unroll
    for (int i = 0; i < N1; ++i) u[i] = T[row * (N1 + 1) + i];
    dft_reg<N1, SIGN>(u);
}

// the same from layout B, u[m] = x[j + N2 m] on lane j < N2, into layout A,
// v[p] = X[j + N1 p] on lane j < N1
template <int N, int SIGN>
__device__ __forceinline__ void zw_b2a(double2 (&u)[ZW<N>::N1], double2 (&v)[ZW<N>::N2], int j, double2* T,
                                       const double2* tw) {
    constexpr int N1 = ZW<N>::N1, N2 = ZW<N>::N2;
    dft_reg<N1, SIGN>(u);
    const int jt = j < N2 ? j : N2 - 1;
#pragma unroll
    for (int q = 1; q < N1; ++q) {
        double2 t = tw[jt * q];
        if (SIGN > 0) t.y = -t.y;
        u[q] = cmul(u[q], t);
    }
    wave_lds_sync();
    if (j < N2) {
#pragma unroll
        for (int q = 0; q < N1; ++q) T[q * (N2 + 1) + j] = u[q];
    }
    wave_lds_sync();
    const int row = j < N1 ? j : N1 - 1;
#pragma unroll
    for (int i = 0; i < N2; ++i) v[i] = T[row * (N2 + 1) + i];
    dft_reg<N2, SIGN>(v);
}

// c2r of two real fields at once (fa -> pa, fb -> pb; fb < 0: none), from the
// live bins kz < nkz of xp; dza / dzb apply d/dz. As in FFTW's c2r, the
// imaginary part of bin 0 is ignored. STAGED: the bins come from staging
// slots 0 (fa) and 1 (fb), and once these are read the copies of fields na and
// nb (< 0: none) into them start, to land while this transform and what
// follows it run.
template <int NZ, bool STAGED>
__device__ __forceinline__ void zw_c2r(const ZWCtx& c, const double2* __restrict__ xp, int fa, bool dza, int fb,
                                       bool dzb, double (&pa)[ZW<NZ>::N1], double (&pb)[ZW<NZ>::N1], int na = -1,
                                       int nb = -1) {
    constexpr int N1 = ZW<NZ>::N1, N2 = ZW<NZ>::N2, N = ZW<NZ>::N;
    const int j = c.gi;
    const bool act = c.line_ok && j < N1;
    const double2* Sa = c.S + c.sline;
    const double2* Sb = Sa + c.sstride;
    double2 v[N2];
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int n = j + N1 * r;
        const bool lo = n < c.nkz, hi = n > N - c.nkz;
        const int k = lo ? n : N - n;
        double2 A = make_double2(0.0, 0.0), B = make_double2(0.0, 0.0);
        if (act && (lo || hi)) {
            if constexpr (STAGED) {
                A = Sa[k];
                if (fb >= 0) B = Sb[k];
            } else {
                A = xp[(long)fa * c.fsz + c.base + k];
                if (fb >= 0) B = xp[(long)fb * c.fsz + c.base + k];
            }
        }
        if (dza) A = imul(c.dk * k, A);
        if (dzb) B = imul(c.dk * k, B);
        if (n == 0) v[r] = make_double2(A.x, B.x);
        else if (lo) v[r] = make_double2(A.x - B.y, A.y + B.x);  // A + iB
        else v[r] = make_double2(A.x + B.y, B.x - A.y);          // conj(A) + i conj(B) (zero in the gap)
    }
    if (STAGED && na >= 0) {
        // every lane's reads of the slots have returned before the copies may overwrite them
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        zw_stage<NZ>(c, xp, na, 0);
        if (nb >= 0) zw_stage<NZ>(c, xp, nb, 1);
    }
    double2 u[N1];
    zw_a2b<NZ, +1>(v, u, j, c.T, c.tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
        pa[m] = u[m].x;
        pb[m] = u[m].y;
    }
}

// r2c of pa (and pb when TWO) -> zo fields ga (and gb), bins kz < nkz.
// fold != 0 (with TWO): field ga receives A + fold i kz B instead of A, with
// kz = 2 pi k / Lz, A and B the spectra of pa and pb; gb < 0 then drops B.
template <int NZ, bool TWO>
__device__ __forceinline__ void zw_r2c(const ZWCtx& c, double2* __restrict__ zo, const double (&pa)[ZW<NZ>::N1],
                                       const double (&pb)[ZW<NZ>::N1], int ga, int gb, double fold = 0.0) {
    constexpr int N1 = ZW<NZ>::N1, N2 = ZW<NZ>::N2, N = ZW<NZ>::N;
    const int j = c.gi;
    double2 v[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m) v[m] = make_double2(pa[m], TWO ? pb[m] : 0.0);
    double2 w[N2];
    zw_b2a<NZ, -1>(v, w, j, c.T, c.tw);  // w[p] = X[j + N1 p]
    wave_lds_sync();
    if (j < N1) {
#pragma unroll
        for (int p = 0; p < N2; ++p) c.T[j + N1 * p] = w[p];
    }
    wave_lds_sync();
    if (c.line_ok && j < N1) {
#pragma unroll
        for (int p = 0; p < N2; ++p) {
            const int k = j + N1 * p;
            if (k < c.nkz) {
                const double2 Zk = w[p];
                const double2 Zn = c.T[(N - k) % N];
                const double2 A = make_double2(0.5 * (Zk.x + Zn.x), 0.5 * (Zk.y - Zn.y));
                const double2 B = make_double2(0.5 * (Zk.y + Zn.y), -0.5 * (Zk.x - Zn.x));
                if (TWO && fold != 0.0) {
                    const double fk = fold * (c.dk * k);
                    zo[(long)ga * c.fsz + c.base + k] = make_double2(A.x - fk * B.y, A.y + fk * B.x);
                } else {
                    zo[(long)ga * c.fsz + c.base + k] = A;
                }
                if (TWO && gb >= 0) zo[(long)gb * c.fsz + c.base + k] = B;
            }
        }
    }
}

// wavefronts per block of the staged wavefront z pass
constexpr int ZWAVE_WB = 1;

// The z pass stages its inputs at Nz = 64, and at Nz = 128 for the rotational
// form. The choice depends on the grid and form alone (never on the member
// count), so a member's result is the same alone and in a batch: the two forms
// contract multiply-adds differently (the compiler fuses either product of
// a*b + c*d depending on the code around it) and differ in the last bits. At
// Nz = 64 the staged form wins where it matters, in batches (16 members of
// 64x65x64, skew-symmetric: 104 against 122 us per launch) and loses 2.6 us on
// a lone field; at 32 and 48 (24 members) and for lone 96^3 and 192^3 fields
// the unstaged one is faster. The rotational form at Nz = 128 runs the staged
// kernel at 2 waves per SIMD (zwave_min_waves; 256 VGPRs and a 60-byte spill,
// where the unstaged one needs 277 + 21 and runs 1): a lone 128^3 field 56.4
// -> 47.5 us. The same cap on the skew-symmetric form spills 232-692 bytes at
// 96-128 and on the rotational one 200 bytes at 192, and those were slower.
constexpr bool zwave_staged(int method, int nz) { return nz == 64 || (nz == 128 && method == NL_ROT); }
constexpr int zwave_min_waves(int method, int nz, bool staged) {
    return (staged && nz == 128 && method == NL_ROT) ? 2 : 1;
}

// The z pass of k_zpass with one wavefront per 64/G lines (same fields, same
// products in the same order). A block is independent wavefronts: any number
// up to 4 (LDS sized at launch), or, STAGED, ZWAVE_WB of them, each staging
// the next transform's two input fields in LDS while it works on the current
// one. Staging hides the loads' latency where many wavefronts share a CU;
// with few wavefronts per SIMD the extra LDS round trip costs more than it
// hides (zwave_staged picks the form per grid and nonlinear form).
template <int METHOD, int NZ, bool STAGED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(zwave_min_waves(METHOD, NZ, STAGED), 8)))
k_zwave(Geom g, const double2* __restrict__ twg, const NLMember* __restrict__ tab,
                                               int want_g, const double2* __restrict__ xp0, double2* __restrict__ zo0,
                                               long sx) {
    using S = ZW<NZ>;
    constexpr int N1 = S::N1, SLOT = S::LPW * (NZ / 2);
    double2 *tw_l, *T_l, *in_l = nullptr;
    if constexpr (STAGED) {
        // Separate LDS arrays (not one dynamic block) let the compiler tell the
        // staging copies apart from the transposes: an LDS read waits for the
        // copies in flight only when it reads the staging array.
        __shared__ double2 tw_s[NZ];
        __shared__ double2 T_s[ZWAVE_WB * S::LPW * S::LBUF];
        __shared__ double2 in_s[ZWAVE_WB * 2 * SLOT];
        tw_l = tw_s;
        T_l = T_s;
        in_l = in_s;
    } else {
        extern __shared__ double2 lds[];
        tw_l = lds;
        T_l = lds + NZ;
    }
    for (int i = threadIdx.x; i < NZ; i += blockDim.x) tw_l[i] = twg[i];
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const int nxc = (g.Nx + S::LPW - 1) / S::LPW;
    const int wg = blockIdx.x * (blockDim.x / 64) + wave;
    const int y = wg / nxc;
    if (y >= g.Ny) return;
    const int L = lane / S::G;
    const int x0 = (wg - y * nxc) * S::LPW, x = x0 + L;
    ZWCtx c;
    c.gi = lane % S::G;
    c.line_ok = x < g.Nx;
    c.base = ((long)y * g.Nx + (c.line_ok ? x : 0)) * g.nkz;
    c.fsz = (long)g.Ny * g.Nx * g.nkz;
    c.nkz = g.nkz;
    c.dk = 2 * PI / g.Lz;
    c.T = T_l + (wave * S::LPW + L) * S::LBUF;
    c.tw = tw_l;
    c.wbase = ((long)y * g.Nx + x0) * g.nkz;
    c.wcount = min(S::LPW, g.Nx - x0) * g.nkz;
    c.lane = lane;
    c.sline = L * g.nkz;
    c.sstride = SLOT;
    c.S = in_l + wave * 2 * SLOT;
    const double rotation = tab[blockIdx.y].rotation;
    const double2* __restrict__ xp = xp0 + blockIdx.y * sx;
    double2* __restrict__ zo = zo0 + blockIdx.y * sx;
    if constexpr (STAGED) {
        zw_stage<NZ>(c, xp, 0, 0);  // u, v: the first transform's pair
        zw_stage<NZ>(c, xp, 1, 1);
    }
    if constexpr (METHOD == NL_ROT) {
        double u0[N1], u1[N1], u2[N1], w0[N1], w1[N1], w2[N1];
        zw_c2r<NZ, STAGED>(c, xp, 0, false, 1, false, u0, u1, 2, 3);
        zw_c2r<NZ, STAGED>(c, xp, 2, false, 3, false, u2, w0, 4, 5);
        zw_c2r<NZ, STAGED>(c, xp, 4, false, 5, false, w1, w2);
        double f0[N1], f1[N1], f2[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) {  // cross(vort, u) (diffops.cpp:2598-2605)
            const double a0 = u0[m], a1 = u1[m], a2 = u2[m], b0 = w0[m], b1 = w1[m], b2 = w2[m];
            f0[m] = b1 * a2 - b2 * a1;
            f1[m] = b2 * a0 - b0 * a2;
            f2[m] = b0 * a1 - b1 * a0;
            if (rotation != 0.0) {  // nse.cpp:61-80
                f0[m] -= rotation * a1;
                f1[m] += rotation * a0;
            }
        }
        zw_r2c<NZ, true>(c, zo, f0, f1, 0, 1);
        zw_r2c<NZ, false>(c, zo, f2, f2, 2, -1);
    } else if constexpr (METHOD == NL_CONV || METHOD == NL_SKEW) {
        // f_i = sum_j (s u_j) du_i/dx_j, terms added in the order j = x, y, z
        // (diffops.cpp:3192-3199, 3626-3633). The skew-symmetric form keeps
        // h = u/2 only: halving is exact, so s u_j is h_j, u_j is 2 h_j and
        // u_i u_j is 4 (h_i h_j), all with the bits of the unscaled products.
        //
        // Skew-symmetric outputs: the z-derivative part of div(u u)/2
        // (diffops.cpp:3254-3279) is added here, in z-spectral space:
        // field i = f_i + 1/2 i kz (u_i u_z) (i = 0, 1, 2), then u_0u_0, u_0u_1,
        // u_0u_2, u_1u_1, u_1u_2 in fields 3-7; u_2u_2 is needed only for that
        // z-derivative and never leaves. The forward transforms add the x and y
        // parts of the divergence.
        constexpr bool SK = (METHOD == NL_SKEW);
        constexpr double s = SK ? 0.5 : 1.0, us = SK ? 2.0 : 1.0;
        double h0[N1], h1[N1], h2[N1], pa[N1], pb[N1], fa[N1], fb[N1];
        zw_c2r<NZ, STAGED>(c, xp, 0, false, 1, false, h0, h1, 2, 6);   // u, v
        zw_c2r<NZ, STAGED>(c, xp, 2, false, 6, false, h2, pa, 3, 0);   // w, du/dx
        if constexpr (SK) {
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                h0[m] *= s;
                h1[m] *= s;
                h2[m] *= s;
            }
        }
#pragma unroll
        for (int m = 0; m < N1; ++m) fa[m] = 0.0 + h0[m] * pa[m];
        zw_c2r<NZ, STAGED>(c, xp, 3, false, 0, true, pa, pb, 7, 4);    // du/dy, du/dz
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            fa[m] += h1[m] * pa[m];
            fa[m] += h2[m] * pb[m];
            if (rotation != 0.0) fa[m] -= rotation * (us * h1[m]);  // nse.cpp:61-80
            if (SK) pb[m] = 4.0 * (h0[m] * h2[m]);
        }
        if constexpr (SK) zw_r2c<NZ, true>(c, zo, fa, pb, 0, 5, 0.5);  // f0 + i kz u0u2 / 2, and u0u2
        else zw_r2c<NZ, false>(c, zo, fa, pb, 0, -1);                  // f0
        zw_c2r<NZ, STAGED>(c, xp, 7, false, 4, false, pa, pb, 1, 8);   // dv/dx, dv/dy
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            fa[m] = 0.0 + h0[m] * pa[m];
            fa[m] += h1[m] * pb[m];
        }
        zw_c2r<NZ, STAGED>(c, xp, 1, true, 8, false, pa, pb, 5, 2);    // dv/dz, dw/dx
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            fa[m] += h2[m] * pa[m];
            if (rotation != 0.0) fa[m] += rotation * (us * h0[m]);
            fb[m] = 0.0 + h0[m] * pb[m];
        }
        zw_c2r<NZ, STAGED>(c, xp, 5, false, 2, true, pa, pb);    // dw/dy, dw/dz
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            fb[m] += h1[m] * pa[m];
            fb[m] += h2[m] * pb[m];
        }
        if constexpr (SK) {
#pragma unroll
            for (int m = 0; m < N1; ++m) pa[m] = 4.0 * (h1[m] * h2[m]);
            zw_r2c<NZ, true>(c, zo, fa, pa, 1, 7, 0.5);   // f1 + i kz u1u2 / 2, and u1u2
#pragma unroll
            for (int m = 0; m < N1; ++m) pa[m] = 4.0 * (h2[m] * h2[m]);
            zw_r2c<NZ, true>(c, zo, fb, pa, 2, -1, 0.5);  // f2 + i kz u2u2 / 2
            // remaining entries of outer(u, u) (diffops.cpp:2336-2388)
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                pa[m] = 4.0 * (h0[m] * h0[m]);
                pb[m] = 4.0 * (h0[m] * h1[m]);
            }
            zw_r2c<NZ, true>(c, zo, pa, pb, 3, 4);
#pragma unroll
            for (int m = 0; m < N1; ++m) pa[m] = 4.0 * (h1[m] * h1[m]);
            zw_r2c<NZ, false>(c, zo, pa, pa, 6, -1);
        } else {
            zw_r2c<NZ, true>(c, zo, fa, fb, 1, 2);            // f1, f2
        }
    } else {
        double u0[N1], u1[N1], u2[N1], pa[N1], pb[N1];
        zw_c2r<NZ, STAGED>(c, xp, 0, false, 1, false, u0, u1, 2);
        zw_c2r<NZ, STAGED>(c, xp, 2, false, -1, false, u2, pa);
        if (want_g) {  // DIV with rotation: g = rotation terms only
#pragma unroll
            for (int m = 0; m < N1; ++m) {
                pa[m] = -(rotation * u1[m]);
                pb[m] = rotation * u0[m];
            }
            zw_r2c<NZ, true>(c, zo, pa, pb, 0, 1);
#pragma unroll
            for (int m = 0; m < N1; ++m) pa[m] = 0.0;
            zw_r2c<NZ, false>(c, zo, pa, pa, 2, -1);
        }
        // outer(u, u) (diffops.cpp:2336-2388), two entries per r2c
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            pa[m] = u0[m] * u0[m];
            pb[m] = u0[m] * u1[m];
        }
        zw_r2c<NZ, true>(c, zo, pa, pb, 3, 4);
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            pa[m] = u0[m] * u2[m];
            pb[m] = u1[m] * u1[m];
        }
        zw_r2c<NZ, true>(c, zo, pa, pb, 5, 6);
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            pa[m] = u1[m] * u2[m];
            pb[m] = u2[m] * u2[m];
        }
        zw_r2c<NZ, true>(c, zo, pa, pb, 7, 8);
    }
}

// ---- wave-local x FFTs ---------------------------------------------------------
// One wavefront per 64/G consecutive kz-lines of one (field, y); each line's
// FFT as in the wave-local z pass. Consecutive lanes of a group hold
// consecutive kx (contiguous reads of yp / writes of fo); the lines of a
// wavefront are consecutive kz (128-byte runs of xp / zo).

// inverse x-FFT: yp [F][Ny][nkz][nkx] -> xp [Fo][Ny][Nx][nkz]; output f >= nsrc
// is d/dx of source field f - nsrc
template <int NX>
__global__ void __launch_bounds__(256) k_xwinv(Geom g, const double2* __restrict__ twg, int nsrc,
                                               const double2* __restrict__ yp0, long sy, double2* __restrict__ xp0,
                                               long sx) {
    const BlockId bid = xcd_grid();  // neighbouring blocks on one XCD
    using S = ZW<NX>;
    constexpr int N1 = S::N1, N2 = S::N2;
    extern __shared__ double2 lds[];
    for (int i = threadIdx.x; i < NX; i += blockDim.x) lds[i] = twg[i];
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nkz = g.nkz, nkx = g.nkx, Ny = g.Ny;
    const int nzc = (nkz + S::LPW - 1) / S::LPW;
    const int wg = bid.x * (blockDim.x / 64) + wave;
    const int y = wg / nzc;
    if (y >= Ny) return;
    const int L = lane / S::G, j = lane % S::G;
    const int ikz = (wg - y * nzc) * S::LPW + L;
    const bool line_ok = ikz < nkz;
    const int fo = bid.y;
    const bool deriv = fo >= nsrc;
    const int fs = deriv ? fo - nsrc : fo;
    const double2* __restrict__ src = yp0 + bid.z * sy + (((long)fs * Ny + y) * nkz + (line_ok ? ikz : 0)) * nkx;
    const double dk = 2 * PI / g.Lx;
    // All of a lane's loads are issued before any is used (bins it does not
    // hold read bin 0 and are zeroed after): a load under the lane condition
    // whose value was scaled at once got its own wait, one HBM round trip per
    // load.
    double2 v[N2];
    int ikr[N2];
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int n = j + N1 * r;
        int ikx = -1;
        if (n <= g.kxd) ikx = n;
        else if (n >= NX - g.kxd) ikx = n - NX + nkx;
        ikr[r] = (line_ok && j < N1) ? ikx : -1;
        v[r] = src[ikr[r] >= 0 ? ikr[r] : 0];
    }
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        double2 z = make_double2(0.0, 0.0);
        if (ikr[r] >= 0) {
            z = v[r];
            if (deriv) z = imul(dk * kx_of(g, ikr[r]), z);
        }
        v[r] = z;
    }
    double2 u[N1];
    zw_a2b<NX, +1>(v, u, j, lds + NX + (wave * S::LPW + L) * S::LBUF, lds);
    if (line_ok && j < N2) {
        double2* __restrict__ out = xp0 + bid.z * sx + ((long)fo * Ny + y) * NX * (long)nkz + ikz;
#pragma unroll
        for (int m = 0; m < N1; ++m) out[(long)(j + N2 * m) * nkz] = u[m];
    }
}

// forward x-FFT: zo [G][Ny][Nx][nkz] -> fo [G][Ny][nkz][nkx], scaled by 1/(Nx Nz).
// fold (the skew-symmetric form): the x-derivative part of its divergence is
// added here, in x-spectral space, and the launch has 5 line tasks, not 8:
// task c < 3 writes field c as f_c + 1/2 i kx (u_c u_0) (u_c u_0 is z-pass
// field 3 + c), task 1 also u_0u_1 as field 3, tasks 3 and 4 u_1u_1 (z-pass
// field 6) and u_1u_2 (7) as fields 4 and 5. The y-forward transform then
// reads f_c and u_c u_1 only. The added term is formed exactly as the
// y-forward transform formed it before (the same products in the same order).
// xcd: blocks in XCD order (xcd_grid), for grids whose rows of nkz bins are
// not a whole number of 128-byte lines, so that neighbouring kz chunks, which
// share the lines at their edges, read them through one L2; where the rows
// are whole lines the dispatch order is faster.
template <int NX>
__global__ void __launch_bounds__(256) k_xwfwd(Geom g, const double2* __restrict__ twg,
                                               const double2* __restrict__ zo0, long sx, double2* __restrict__ fo0,
                                               long sy, int fold, int xcd) {
    const BlockId bid = xcd ? xcd_grid() : BlockId{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
    using S = ZW<NX>;
    constexpr int N1 = S::N1, N2 = S::N2;
    extern __shared__ double2 lds[];
    for (int i = threadIdx.x; i < NX; i += blockDim.x) lds[i] = twg[i];
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nkz = g.nkz, nkx = g.nkx, Ny = g.Ny;
    const int nzc = (nkz + S::LPW - 1) / S::LPW;
    const int wg = bid.x * (blockDim.x / 64) + wave;
    const int y = wg / nzc;
    if (y >= Ny) return;
    const int L = lane / S::G, j = lane % S::G;
    const int ikz = (wg - y * nzc) * S::LPW + L;
    const bool line_ok = ikz < nkz;
    const int t = bid.y;
    int fa = t, oa = t, fb = -1, ob = -1;
    if (fold) {
        fa = t < 3 ? t : t + 3;
        oa = t < 3 ? t : t + 1;
        if (t < 3) fb = 3 + t;
        if (t == 1) ob = 3;
    }
    double2* const T = lds + NX + (wave * S::LPW + L) * S::LBUF;
    auto line = [&](int f, double2 (&v)[N2]) {
        const double2* __restrict__ in =
            zo0 + bid.z * sx + ((long)f * Ny + y) * NX * (long)nkz + (line_ok ? ikz : 0);
        double2 u[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m)
            u[m] = (line_ok && j < N2) ? in[(long)(j + N2 * m) * nkz] : make_double2(0.0, 0.0);
        zw_b2a<NX, -1>(u, v, j, T, lds);
    };
    double2 va[N2], vb[N2];
    line(fa, va);
    if (fb >= 0) line(fb, vb);
    if (line_ok && j < N1) {
        const double scale = 1.0 / ((double)g.Nx * g.Nz), beta = 0.5;
        double2* __restrict__ base = fo0 + bid.z * sy;
        double2* __restrict__ outa = base + (((long)oa * Ny + y) * nkz + ikz) * nkx;
        double2* __restrict__ outb = base + (((long)(ob < 0 ? oa : ob) * Ny + y) * nkz + ikz) * nkx;
#pragma unroll
        for (int p = 0; p < N2; ++p) {
            const int n = j + N1 * p;
            int ikx = -1;
            if (n <= g.kxd) ikx = n;
            else if (n >= NX - g.kxd) ikx = n - NX + nkx;
            if (ikx < 0) continue;
            double2 a = make_double2(va[p].x * scale, va[p].y * scale);
            if (fb >= 0) {
                const double2 b = make_double2(vb[p].x * scale, vb[p].y * scale);
                const int kx = kx_of(g, ikx);
                const double s0 = -beta * 2 * PI * kx / g.Lx, s1 = beta * 2 * PI * kx / g.Lx;
                a = make_double2(fma(s0, b.y, a.x), fma(s1, b.x, a.y));  // a + beta i kx b
                if (ob >= 0) outb[ikx] = b;
            }
            outa[ikx] = a;
        }
    }
}

// ---- y transforms as LDS-tiled contractions ----------------------------------
// Both directions are dense Ny x Ny products over the Chebyshev index applied to
// every (component, mode) column. A block computes YB_M modes x YB_R output
// rows of one member; the contraction index runs through LDS in chunks of
// YB_K, and each thread accumulates a 2 x 2 (mode, row) micro-tile of every
// product it needs, so the epilogue that mixes components (vorticity, base
// flow, divergence terms) stays in registers. Row tiles of one mode tile are
// adjacent in launch order, so the input tile's re-reads come from L2.
constexpr int YB_M = 32, YB_R = 32, YB_K = 16;

// Chebyshev -> y values (same outputs as k_yinv). QM: which components need T*D.
template <int METHOD>
__global__ void __launch_bounds__(256) k_yinv_t(Geom g, YMats ym, const NLMember* __restrict__ tab,
                                                double2* __restrict__ yp0, long sy) {
    constexpr bool Q0 = (METHOD != NL_DIV), Q1 = (METHOD == NL_CONV || METHOD == NL_SKEW), Q2 = (METHOD != NL_DIV);
    constexpr bool QA = Q0 || Q1 || Q2;
    __shared__ double2 Xs[3][YB_K][YB_M];
    __shared__ double As[2][YB_R][YB_K + 1];
    const NLMember& me = tab[blockIdx.y];
    const double2* __restrict__ u = reinterpret_cast<const double2*>(me.u);
    double2* __restrict__ yp = yp0 + blockIdx.y * sy;
    const int Ny = g.Ny;
    const long M = g.M;
    const int nrt = (Ny + YB_R - 1) / YB_R;
    const int m0 = (blockIdx.x / nrt) * YB_M, y0 = (blockIdx.x % nrt) * YB_R;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double2 P[3][2][2], Q[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) P[c][i][j] = Q[c][i][j] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < Ny; k0 += YB_K) {
        for (int i = threadIdx.x; i < 3 * YB_K * YB_M; i += 256) {
            const int mm = i % YB_M, kk = (i / YB_M) % YB_K, c = i / (YB_M * YB_K);
            const int k = k0 + kk, m = m0 + mm;
            Xs[c][kk][mm] = (k < Ny && m < M) ? u[((long)c * Ny + k) * M + m] : make_double2(0.0, 0.0);
        }
        for (int i = threadIdx.x; i < (QA ? 2 : 1) * YB_R * YB_K; i += 256) {
            const int kk = i % YB_K, r = (i / YB_K) % YB_R, a = i / (YB_K * YB_R);
            const int k = k0 + kk, y = y0 + r;
            As[a][r][kk] = (k < Ny && y < Ny) ? (a == 0 ? ym.T : ym.TD)[(long)y * Ny + k] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < YB_K; ++kk) {
            double t[2], d[2];
            double2 x[3][2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                t[j] = As[0][ty + 16 * j][kk];
                d[j] = QA ? As[1][ty + 16 * j][kk] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int i = 0; i < 2; ++i) x[c][i] = Xs[c][kk][tx + 16 * i];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        P[c][i][j].x += t[j] * x[c][i].x;
                        P[c][i][j].y += t[j] * x[c][i].y;
                    }
                    if (Q0) { Q[0][i][j].x += d[j] * x[0][i].x; Q[0][i][j].y += d[j] * x[0][i].y; }
                    if (Q1) { Q[1][i][j].x += d[j] * x[1][i].x; Q[1][i][j].y += d[j] * x[1][i].y; }
                    if (Q2) { Q[2][i][j].x += d[j] * x[2][i].x; Q[2][i][j].y += d[j] * x[2][i].y; }
                }
        }
        __syncthreads();
    }
    const double* __restrict__ base = me.base;
    const double Vsuck = me.Vsuck;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = m0 + tx + 16 * i;
        if (m >= M) continue;
        const int ikx = m % g.nkx, ikz = m / g.nkx;
        const double dx = 2 * PI * kx_of(g, ikx) / g.Lx, dz = 2 * PI * ikz / g.Lz;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int y = y0 + ty + 16 * j;
            if (y >= Ny) continue;
            double2 p0 = P[0][i][j], p1 = P[1][i][j], p2 = P[2][i][j];
            double2 q0 = Q[0][i][j], q1 = Q[1][i][j], q2 = Q[2][i][j];
            if (m == 0) {  // base flow on the kx = kz = 0 mode (nse.cpp:28-36)
                p0.x += base[0 * Ny + y];
                q0.x += base[1 * Ny + y];
                p2.x += base[2 * Ny + y];
                q2.x += base[3 * Ny + y];
                p1.x -= Vsuck;
            }
            double2 o[6];
            o[0] = p0;
            o[1] = p1;
            o[2] = p2;
            int nout = 3;
            if (METHOD == NL_ROT) {
                o[3] = csub(q2, imul(dz, p1));               // dw/dy - dv/dz
                o[4] = csub(imul(dz, p0), imul(dx, p2));     // du/dz - dw/dx
                o[5] = csub(imul(dx, p1), q0);               // dv/dx - du/dy
                nout = 6;
            } else if (METHOD == NL_CONV || METHOD == NL_SKEW) {
                o[3] = q0;
                o[4] = q1;
                o[5] = q2;
                nout = 6;
            }
            for (int f = 0; f < nout; ++f) yp[((long)f * Ny + y) * M + m] = o[f];
        }
    }
}

// y values -> Chebyshev coefficients (same outputs as k_yfwd): f_i = Ti G_i + DTi H_i.
template <int METHOD>
__global__ void __launch_bounds__(256) k_yfwd_t(Geom g, YMats ym, int have_g, const double2* __restrict__ fo0,
                                                long sy, const NLMember* __restrict__ tab, int xfold) {
    constexpr bool DIVF = (METHOD == NL_DIV || METHOD == NL_SKEW);
    constexpr int NX = DIVF ? 6 : 3;
    __shared__ double2 Xs[NX][YB_K][YB_M];
    __shared__ double As[2][YB_R][YB_K + 1];
    const double2* __restrict__ fo = fo0 + blockIdx.y * sy;
    double2* __restrict__ f = reinterpret_cast<double2*>(tab[blockIdx.y].f);
    const int Ny = g.Ny;
    const long M = g.M;
    const int nrt = (Ny + YB_R - 1) / YB_R;
    const int m0 = (blockIdx.x / nrt) * YB_M, k0o = (blockIdx.x % nrt) * YB_R;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const double beta = (METHOD == NL_SKEW) ? 0.5 : 1.0;
    double2 Acc[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) Acc[c][i][j] = make_double2(0.0, 0.0);
    for (int y0 = 0; y0 < Ny; y0 += YB_K) {
        // G_i = g_i + beta (ikx uu_i0 + ikz uu_i2), H_i = beta uu_i1 (uu index -> fo field: 00:3 01:4 02:5 11:6
        // 12:7 22:8). The skew-symmetric z pass has already added beta ikz uu_i2 to g_i (and keeps no uu22).
        for (int i = threadIdx.x; i < 3 * YB_K * YB_M; i += 256) {
            const int mm = i % YB_M, kk = (i / YB_M) % YB_K, c = i / (YB_M * YB_K);
            const int y = y0 + kk, m = m0 + mm;
            double2 G = make_double2(0.0, 0.0), H = make_double2(0.0, 0.0);
            if (y < Ny && m < M) {
                if (have_g) G = fo[((long)c * Ny + y) * M + m];
                if (DIVF && METHOD == NL_SKEW && xfold) {  // x and z terms already in g_c; uu_c1 is field 3 + c
                    const double2 x1 = fo[((long)(3 + c) * Ny + y) * M + m];
                    H = make_double2(beta * x1.x, beta * x1.y);
                } else if (DIVF) {
                    const int ikx = m % g.nkx, ikz = m / g.nkx;
                    const double dx = 2 * PI * kx_of(g, ikx) / g.Lx, dz = 2 * PI * ikz / g.Lz;
                    const int uuf[3][3] = {{3, 4, 5}, {4, 6, 7}, {5, 7, 8}};
                    const double2 x0 = fo[((long)uuf[c][0] * Ny + y) * M + m];
                    const double2 x1 = fo[((long)uuf[c][1] * Ny + y) * M + m];
                    double2 d = imul(dx, x0);
                    if (METHOD == NL_DIV) d = cadd(d, imul(dz, fo[((long)uuf[c][2] * Ny + y) * M + m]));
                    G = make_double2(G.x + beta * d.x, G.y + beta * d.y);
                    H = make_double2(beta * x1.x, beta * x1.y);
                }
            }
            Xs[c][kk][mm] = G;
            if (DIVF) Xs[3 + c][kk][mm] = H;
        }
        for (int i = threadIdx.x; i < (DIVF ? 2 : 1) * YB_R * YB_K; i += 256) {
            const int kk = i % YB_K, r = (i / YB_K) % YB_R, a = i / (YB_K * YB_R);
            const int y = y0 + kk, k = k0o + r;
            As[a][r][kk] = (k < Ny && y < Ny) ? (a == 0 ? ym.Ti : ym.DTi)[(long)k * Ny + y] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < YB_K; ++kk) {
            double t[2], d[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                t[j] = As[0][ty + 16 * j][kk];
                d[j] = DIVF ? As[1][ty + 16 * j][kk] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double2 xg[2], xh[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    xg[i] = Xs[c][kk][tx + 16 * i];
                    xh[i] = DIVF ? Xs[3 * DIVF + c][kk][tx + 16 * i] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        Acc[c][i][j].x += t[j] * xg[i].x;
                        Acc[c][i][j].y += t[j] * xg[i].y;
                        if (DIVF) {
                            Acc[c][i][j].x += d[j] * xh[i].x;
                            Acc[c][i][j].y += d[j] * xh[i].y;
                        }
                    }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = m0 + tx + 16 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int k = k0o + ty + 16 * j;
            if (k >= Ny) continue;
            for (int c = 0; c < 3; ++c) f[((long)c * Ny + k) * M + m] = Acc[c][i][j];
        }
    }
}

// ---- CFL factor -------------------------------------------------------------
// FlowField::CFLfactor(Ubase, Wbase) as DNSAlgorithm::CFL calls it
// (flowfield.cpp:4035-4068, dnsalgo.cpp:63-67): the largest of
// (u_i + U_i(y)) / dx_i over every grid point and component, signed as in the
// reference, floored at 0; dx_y is the local Chebyshev spacing. The physical
// velocity comes from the divergence form's inverse transforms, which add U, W
// and -Vsuck to the (0,0) mode; Vsuck is added back for v. One partial maximum
// per block, then one per member.
template <int NZ>
__global__ void __launch_bounds__(256) k_zcfl(Geom g, FFTPlan pz, int XB, const NLMember* __restrict__ tab,
                                              const double2* __restrict__ xp0, long sx, double* __restrict__ part) {
    extern __shared__ double2 lds[];
    __shared__ double red[4];
    const double vs = tab[blockIdx.y].Vsuck;
    const double2* __restrict__ xp = xp0 + blockIdx.y * sx;
    const int nxc = (g.Nx + XB - 1) / XB;
    ZCtx c;
    c.g = g;
    c.pz = pz;
    c.y = blockIdx.x / nxc;
    c.x0 = (blockIdx.x % nxc) * XB;
    c.nl = min(XB, g.Nx - c.x0);
    c.XB = XB;
    c.tw = lds;
    c.wa = lds + g.Nz;
    c.wb = c.wa + XB * lpitch(g.Nz);
    c.phys = reinterpret_cast<double*>(c.wb + XB * lpitch(g.Nz));
    stage_tw(c.tw, pz);
    zc2r_pair<NZ>(c, xp, 0, false, 0, 1, false, 1);
    zc2r_pair<NZ>(c, xp, 2, false, 2, -1, false, -1);
    const int Ny = g.Ny, y = c.y;
    const double piN = PI / (Ny - 1), rad = (g.b - g.a) / 2, cen = (g.b + g.a) / 2;  // chebypoints (chebyshev.cpp:746-754)
    auto yc = [&](int j) { return cen + rad * cos(j * piN); };
    const double dy = (y == 0 || y == Ny - 1) ? yc(0) - yc(1) : (yc(y - 1) - yc(y + 1)) / 2.0;
    const double dx = g.Lx / g.Nx, dz = g.Lz / g.Nz;
    const double *u0 = physf(c, 0), *u1 = physf(c, 1), *u2 = physf(c, 2);
    double m = 0.0;
    for (int w = threadIdx.x; w < c.nl * g.Nz; w += blockDim.x) {
        m = fmax(m, u0[w] / dx);
        m = fmax(m, (u1[w] + vs) / dy);
        m = fmax(m, u2[w] / dz);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
        part[(long)blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(256) k_cfl_reduce(int nb, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double red[4];
    const double* p = part + (long)blockIdx.x * nb;
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, p[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
        out[blockIdx.x] = m;
    }
}

// ---- y transforms on the FP64 matrix cores ----------------------------------
// v_mfma_f64_16x16x4f64 (one wave: D = A B + C, A 16x4, B 4x16, D 16x16).
// Operand layout, lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result
// register r of lane l is D[4 r + l / 16][l % 16].
//
// A block is three waves per transform matrix; wave c handles velocity
// component c of YW_MODES = 8 consecutive modes, i.e. 16 real columns (real,
// imaginary part of each mode), over all output rows. Each lane's B operand
// streams from HBM straight into registers, one chunk ahead; the chunk of the
// transform matrices, shared by the waves, is staged in LDS.
//
// Both transforms use the mirror symmetry of the Chebyshev points,
// y_{Ny-1-j} = -y_j: T[Ny-1-j][k] = (-1)^k T[j][k], and the derivative flips
// the parity. Only the upper half of the points (j < (Ny+1)/2) enters the
// products, split by the parity of k, which halves the matrix work.
typedef double mfma4d __attribute__((ext_vector_type(4)));
constexpr int YW_MODES = 8;

__device__ __forceinline__ mfma4d mfma16x16x4(double a, double b, mfma4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Waves per SIMD the register allocation of a wave-column kernel must allow:
// the accumulators take 16 NRT registers, the rest is operand staging (without
// a bound the scheduler hoists a whole chunk's LDS reads) and, in the inverse,
// the next matrix chunk in flight.
constexpr int yw_waves(int nrt) { return nrt <= 3 ? 4 : (nrt <= 4 ? 3 : 2); }
// The forward transform also holds sums and differences of mirrored values of
// the prefetched chunk, and the next matrix chunk in flight (two matrices in
// the divergence forms).
constexpr int yw_waves_fwd(bool divf, int nrt) { return nrt <= 1 ? 4 : (nrt <= (divf ? 2 : 3) ? 3 : 2); }

// Chebyshev -> y values (outputs as k_yinv_t): p = T u and, unless the form is
// the divergence one, q = T D u = du/dy values. Waves 0-2 compute p of
// components 0-2, waves 3-5 q; the two halves read the same coefficient
// columns (the second read is served by the caches). For an upper point
// j < Ny/2, E = sum over even k, O = sum over odd k: p_j = E + O,
// p_{Ny-1-j} = E - O, q_j = E + O, q_{Ny-1-j} = O - E. The middle point of an
// odd Ny is a dot product per lane, summed over the lane groups. NRT: row
// tiles of 16 upper points.
template <int METHOD, int NRT>
__global__ void __launch_bounds__(METHOD == NL_DIV ? 192 : 384) __attribute__((amdgpu_waves_per_eu(yw_waves(NRT))))
    k_yinv_w(Geom g, YMats ym, const NLMember* __restrict__ tab, double2* __restrict__ yp0, long sy) {
    const BlockId bid = xcd_grid();  // neighbouring blocks on one XCD
    constexpr int NMAT = (METHOD == NL_DIV) ? 1 : 2, NT = 192 * NMAT;
    constexpr int KC = 16, AP = KC + 1, EP = 17;  // chunk columns: 8 even k, then 8 odd k
    constexpr int RS = NRT * 16 + 1;              // LDS rows per matrix: tiles, middle point
    // two matrix chunk buffers: chunk i + 1 is loaded into registers while
    // chunk i is multiplied, and stored to the other buffer after it
    constexpr int SA = NMAT * RS * AP, SE = NMAT * 3 * 2 * 16 * EP;
    constexpr int NSTG = NMAT * RS * KC, NST = (NSTG + NT - 1) / NT;  // staged entries per chunk, per thread
    __shared__ double sm[2 * SA > SE ? 2 * SA : SE];
    const NLMember& me = tab[bid.y];
    double2* __restrict__ yp = yp0 + bid.y * sy;
    const int Ny = g.Ny, nm = Ny / 2, mid = (Ny & 1) ? nm : -1;
    const long M = g.M;
    const int nrt = (nm + 15) / 16;
    const int w = threadIdx.x >> 6, c = w % 3, am = w / 3, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    const long m0 = (long)bid.x * YW_MODES;
    const long mcol = m0 + (lr >> 1);
    const bool mok = mcol < M;
    const gcdouble ucol = gmem(me.u) + 2 * ((long)c * Ny * M + mcol) + (l & 1);
    mfma4d ae[NRT], ao[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) ae[rt] = ao[rt] = mfma4d{0.0, 0.0, 0.0, 0.0};
    double dmid = 0.0;
    double bn[4];  // [parity * 2 + step]: k = k0 + 2 (4 step + lane / 16) + parity
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + 2 * (4 * (i & 1) + lk) + (i >> 1);
            bn[i] = (mok && k < Ny) ? ucol[2L * k * M] : 0.0;
        }
    };
    // matrix entries of chunk k0: entry i = (matrix a, row slot, column kk),
    // stored with the 8 even columns before the 8 odd ones
    double mv[NST];
    auto load_mat = [&](int k0) {
#pragma unroll
        for (int s = 0; s < NST; ++s) {
            const int i = threadIdx.x + s * NT;
            const int kk = i % KC, slot = (i / KC) % RS, a = i / (KC * RS);
            const int k = k0 + kk, row = slot < NRT * 16 ? slot : mid;
            const bool ok = i < NSTG && (slot < NRT * 16 ? slot < nm : mid >= 0) && k < Ny;
            mv[s] = ok ? (a == 0 ? ym.T : ym.TD)[(long)row * Ny + k] : 0.0;
        }
    };
    fetch(0);
    load_mat(0);
    for (int k0 = 0, buf = 0; k0 < Ny; k0 += KC, buf ^= 1) {
        double b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = bn[i];
        // buffer buf was last read two chunks ago, before every thread passed
        // the previous chunk's barrier
        double* __restrict__ sb = sm + buf * SA;
#pragma unroll
        for (int s = 0; s < NST; ++s) {
            const int i = threadIdx.x + s * NT;
            const int kk = i % KC, slot = (i / KC) % RS, a = i / (KC * RS);
            if (i < NSTG) sb[(a * RS + slot) * AP + (kk & 1) * 8 + (kk >> 1)] = mv[s];
        }
        __syncthreads();
        if (k0 + KC < Ny) {
            fetch(k0 + KC);
            load_mat(k0 + KC);
        }
        const double* __restrict__ arow = sb + (am * RS + lr) * AP + lk;
        const double* __restrict__ amid = sb + (am * RS + NRT * 16) * AP + lk;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            // unconditional: staged rows and columns past the grid are zero,
            // and a conditional update keeps the tiles from staying in place
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) {
                ae[rt] = mfma16x16x4(arow[rt * 16 * AP + 4 * st], b[st], ae[rt]);
                ao[rt] = mfma16x16x4(arow[rt * 16 * AP + 8 + 4 * st], b[2 + st], ao[rt]);
            }
            dmid = fma(amid[4 * st], b[st], fma(amid[8 + 4 * st], b[2 + st], dmid));
            __builtin_amdgcn_sched_barrier(0);  // keep each step's operand reads next to its MFMAs
        }
    }
    const int pm = threadIdx.x & 7;
    const long m = m0 + pm;
    const int ikx = (int)(m % g.nkx), ikz = (int)(m / g.nkx);
    const double dx = 2 * PI * kx_of(g, ikx) / g.Lx, dz = 2 * PI * ikz / g.Lz;
    // outputs of one point from the exchange rows: p of component cc at e + cc * ps, q at e + (3 + cc) * ps
    auto emit = [&](int y, const double* e, int ps) {
        double2 p[3], q[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            p[cc] = make_double2(e[cc * ps], e[cc * ps + 1]);
            q[cc] = NMAT == 2 ? make_double2(e[(3 + cc) * ps], e[(3 + cc) * ps + 1]) : make_double2(0.0, 0.0);
        }
        if (m == 0) {  // base flow on the kx = kz = 0 mode (nse.cpp:28-36)
            const gcdouble base = gmem(me.base);
            p[0].x += base[0 * Ny + y];
            q[0].x += base[1 * Ny + y];
            p[2].x += base[2 * Ny + y];
            q[2].x += base[3 * Ny + y];
            p[1].x -= me.Vsuck;
        }
        double2 o[6];
        o[0] = p[0];
        o[1] = p[1];
        o[2] = p[2];
        int nout = 3;
        if (METHOD == NL_ROT) {
            o[3] = csub(q[2], imul(dz, p[1]));            // dw/dy - dv/dz
            o[4] = csub(imul(dz, p[0]), imul(dx, p[2]));  // du/dz - dw/dx
            o[5] = csub(imul(dx, p[1]), q[0]);            // dv/dx - du/dy
            nout = 6;
        } else if (METHOD == NL_CONV || METHOD == NL_SKEW) {
            o[3] = q[0];
            o[4] = q[1];
            o[5] = q[2];
            nout = 6;
        }
        for (int f = 0; f < nout; ++f) yp[((long)f * Ny + y) * M + m] = o[f];
    };
    // row tiles: the components and both halves meet in LDS
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
        if (rt >= nrt) continue;  // uniform
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double e = ae[rt][r], o = ao[rt][r];
            sm[(((am * 3 + c) * 2 + 0) * 16 + 4 * r + lk) * EP + lr] = e + o;
            sm[(((am * 3 + c) * 2 + 1) * 16 + 4 * r + lk) * EP + lr] = am == 0 ? e - o : o - e;
        }
        __syncthreads();
        for (int pt = threadIdx.x; pt < 256; pt += NT) {
            const int half = pt >> 7, pr = (pt >> 3) & 15;
            const int yt = rt * 16 + pr;
            if (yt >= nm || m >= M) continue;
            emit(half ? Ny - 1 - yt : yt, sm + (half * 16 + pr) * EP + 2 * pm, 2 * 16 * EP);
        }
    }
    if (mid < 0) return;  // uniform
    dmid += __shfl_xor(dmid, 16);
    dmid += __shfl_xor(dmid, 32);
    __syncthreads();
    if (lk == 0) sm[(am * 3 + c) * EP + lr] = dmid;
    __syncthreads();
    if (threadIdx.x < YW_MODES && m < M) emit(mid, sm + 2 * pm, EP);
}

// y values -> Chebyshev coefficients (outputs as k_yfwd_t): f_c = Ti G_c +
// DTi H_c, with G_c = g_c + beta (i kx uu_c0 + i kz uu_c2) and H_c = beta uu_c1
// for the divergence forms (beta 1/2 in the skew-symmetric one), G_c = g_c
// otherwise. With S, D the sum and difference of the values at an upper point
// and its mirror (the middle point of an odd Ny counted once):
// f_{2i} = Ti S_G + DTi D_H, f_{2i+1} = Ti D_G + DTi S_H over the upper points.
// The MFMA tiles hold i < Ny/2; the last coefficient of an odd Ny (k = Ny-1)
// is a dot product per lane, summed over the lane groups. NRT: row tiles of 16
// coefficients of each parity.
template <int METHOD, int NRT>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(yw_waves_fwd(METHOD == NL_DIV || METHOD == NL_SKEW, NRT)))) k_yfwd_w(Geom g, YMats ym, int have_g, const double2* __restrict__ fo0,
                                                long sy, const NLMember* __restrict__ tab, int xfold) {
    const BlockId bid = xcd_grid();  // neighbouring blocks on one XCD
    constexpr bool DIVF = (METHOD == NL_DIV || METHOD == NL_SKEW);
    constexpr int NMAT = DIVF ? 2 : 1;
    constexpr int KC = 8, AP = KC + 1, NS = KC / 4;
    constexpr int RS = 2 * NRT * 16 + 1;  // LDS rows per matrix: even, odd coefficients, last one
    // two matrix chunk buffers, as in k_yinv_w, unless the chunk in flight
    // would not fit the registers (two matrices of 6 or 8 row tiles): then one
    // buffer, loaded and stored between two barriers
    constexpr int SA = NMAT * RS * AP, NSTG = NMAT * RS * KC, NST = (NSTG + 191) / 192;
    constexpr bool DB = !(DIVF && NRT >= 6);
    __shared__ double sm[(DB ? 2 : 1) * SA];
    const double* __restrict__ fo = reinterpret_cast<const double*>(fo0 + bid.y * sy);
    __attribute__((address_space(1))) double* __restrict__ f = gmem_w(tab[bid.y].f);
    const int Ny = g.Ny, nh = (Ny + 1) / 2, nm = Ny / 2;
    const bool odd = Ny & 1;
    const long M = g.M;
    const int c = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4, part = l & 1;
    const long mcol = (long)bid.x * YW_MODES + (lr >> 1);
    const bool mok = mcol < M;
    const double beta = (METHOD == NL_SKEW) ? 0.5 : 1.0;
    double sdx = 0, sdz = 0;  // i k multiplication of the other part, with its sign
    if (DIVF) {
        const int ikx = (int)(mcol % g.nkx), ikz = (int)(mcol / g.nkx);
        const double sg = part ? beta : -beta;
        sdx = sg * 2 * PI * kx_of(g, ikx) / g.Lx;
        sdz = sg * 2 * PI * ikz / g.Lz;
    }
    // field of uu_cj in fo: 00:3 01:4 02:5 11:6 12:7 22:8. The skew-symmetric
    // z pass has already added beta ikz uu_c2 to g_c (and keeps no uu22); with
    // xfold the forward x-FFT has added beta ikx uu_c0 too, and uu_c1 is field
    // 3 + c.
    constexpr bool ZTERM = (METHOD == NL_DIV);
    const bool XTERM = !(METHOD == NL_SKEW && xfold);
    const int f0 = c == 0 ? 3 : (c == 1 ? 4 : 5), f2 = c == 0 ? 5 : (c == 1 ? 7 : 8);
    const int f1 = XTERM ? (c == 0 ? 4 : (c == 1 ? 6 : 7)) : 3 + c;
    const long rowst = 2 * M;
    // columns past M read mode M - 1 (every field of fo has Ny x M modes); the
    // values are masked where they are used
    const long mcl = mok ? mcol : M - 1;
    const double* __restrict__ pg = fo + 2 * ((long)c * Ny * M + mcl) + part;
    const double* __restrict__ p0 = fo + 2 * ((long)f0 * Ny * M + mcl) + (1 - part);
    const double* __restrict__ p1 = fo + 2 * ((long)f1 * Ny * M + mcl) + part;
    const double* __restrict__ p2 = fo + 2 * ((long)f2 * Ny * M + mcl) + (1 - part);
    mfma4d ae[NRT], ao[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) ae[rt] = ao[rt] = mfma4d{0.0, 0.0, 0.0, 0.0};
    double dlast = 0.0;
    // Raw values of a chunk, per step: g at the upper point and its mirror,
    // then (divergence forms) uu_c1, then the x and z terms of G. They are
    // loaded without branches, at clamped addresses, and combined only when
    // the chunk is multiplied: a load under a lane condition whose value is
    // used at once gets its own wait, and the chunk's loads were then
    // serialized instead of in flight behind the previous chunk's products.
    constexpr int NR = DIVF ? 8 : 2;
    double rv[NS][NR];
    auto fetch = [&](int y0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int y = min(y0 + 4 * s + lk, nh - 1), ymr = Ny - 1 - y;
            rv[s][0] = pg[y * rowst];
            rv[s][1] = pg[ymr * rowst];
            if (DIVF) {
                rv[s][2] = p1[y * rowst];
                rv[s][3] = p1[ymr * rowst];
                if (XTERM) {
                    rv[s][4] = p0[y * rowst];
                    rv[s][5] = p0[ymr * rowst];
                }
                if (ZTERM) {
                    rv[s][6] = p2[y * rowst];
                    rv[s][7] = p2[ymr * rowst];
                }
            }
        }
    };
    // G at a point from its raw values (i: 0 the upper point, 1 its mirror)
    auto gval = [&](const double* r, int i) {
        double G = have_g ? r[i] : 0.0;
        if (DIVF && XTERM) G += ZTERM ? sdx * r[4 + i] + sdz * r[6 + i] : sdx * r[4 + i];
        return G;
    };
    // matrix entries of chunk y0: entry i = (matrix a, row slot, column kk)
    double mv[NST];
    auto load_mat = [&](int y0) {
#pragma unroll
        for (int s = 0; s < NST; ++s) {
            const int i = threadIdx.x + s * 192;
            const int kk = i % KC, slot = (i / KC) % RS, a = i / (KC * RS);
            const int par = slot / (NRT * 16), r = slot % (NRT * 16), y = y0 + kk;  // par 2: the last coefficient
            const int k = par < 2 ? 2 * r + par : Ny - 1;
            const bool ok = i < NSTG && (par < 2 ? r < nm : odd) && y < nh;
            mv[s] = ok ? (a == 0 ? ym.Ti : ym.DTi)[(long)k * Ny + y] : 0.0;
        }
    };
    fetch(0);
    if (DB) load_mat(0);
    for (int y0 = 0, buf = 0; y0 < nh; y0 += KC, buf ^= DB) {
        double sg[NS], dg[NS], sh[NS], dh[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int y = y0 + 4 * s + lk, ymr = Ny - 1 - y;
            double G = 0, Gm = 0, H = 0, Hm = 0;
            if (mok && y < nh) {
                G = gval(rv[s], 0);
                if (DIVF) H = beta * rv[s][2];
                if (ymr != y) {
                    Gm = gval(rv[s], 1);
                    if (DIVF) Hm = beta * rv[s][3];
                }
            }
            sg[s] = G + Gm;
            dg[s] = G - Gm;
            sh[s] = H + Hm;
            dh[s] = H - Hm;
        }
        // DB: buffer buf was last read two chunks ago, before every thread
        // passed the previous chunk's barrier
        double* __restrict__ sb = sm + buf * SA;
        if (DB) {
#pragma unroll
            for (int s = 0; s < NST; ++s) {
                const int i = threadIdx.x + s * 192;
                const int kk = i % KC, slot = (i / KC) % RS, a = i / (KC * RS);
                if (i < NSTG) sb[(a * RS + slot) * AP + kk] = mv[s];
            }
        } else {
            __syncthreads();  // the previous chunk's matrix reads are done
            for (int i = threadIdx.x; i < NSTG; i += 192) {
                const int kk = i % KC, slot = (i / KC) % RS, a = i / (KC * RS);
                const int par = slot / (NRT * 16), r = slot % (NRT * 16), y = y0 + kk;
                const int k = par < 2 ? 2 * r + par : Ny - 1;
                const bool ok = (par < 2 ? r < nm : odd) && y < nh;
                sb[(a * RS + slot) * AP + kk] = ok ? (a == 0 ? ym.Ti : ym.DTi)[(long)k * Ny + y] : 0.0;
            }
        }
        __syncthreads();
        if (y0 + KC < nh) {
            fetch(y0 + KC);
            if (DB) load_mat(y0 + KC);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            // unconditional: staged rows and columns past the grid are zero,
            // and a conditional update keeps the tiles from staying in place
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) {
                const double* A = sb + (rt * 16 + lr) * AP + 4 * s + lk;
                ae[rt] = mfma16x16x4(A[0], sg[s], ae[rt]);
                ao[rt] = mfma16x16x4(A[NRT * 16 * AP], dg[s], ao[rt]);
                if (DIVF) {
                    ae[rt] = mfma16x16x4(A[RS * AP], dh[s], ae[rt]);
                    ao[rt] = mfma16x16x4(A[(RS + NRT * 16) * AP], sh[s], ao[rt]);
                }
            }
            const double* A = sb + 2 * NRT * 16 * AP + 4 * s + lk;
            dlast = fma(A[0], sg[s], dlast);
            if (DIVF) dlast = fma(A[RS * AP], dh[s], dlast);
            __builtin_amdgcn_sched_barrier(0);  // keep each step's operand reads next to its MFMAs
        }
    }
    dlast += __shfl_xor(dlast, 16);
    dlast += __shfl_xor(dlast, 32);
    if (!mok) return;
    __attribute__((address_space(1))) double* __restrict__ fc = f + 2 * ((long)c * Ny * M + mcol) + part;
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = rt * 16 + 4 * r + lk;
            if (i < nm) {
                fc[2 * i * rowst] = ae[rt][r];
                fc[(2 * i + 1) * rowst] = ao[rt][r];
            }
        }
    }
    if (odd && lk == 0) fc[(Ny - 1) * rowst] = dlast;
}

// Row-tile count classes with a wave-column kernel; above the last one the
// vector tiles run.
template <int METHOD>
static void launch_yinv(const Geom& g, const YMats& ym, int n, const NLMember* tab, double2* yp, long sy,
                        hipStream_t st) {
    const int nrt = (g.Ny / 2 + 15) / 16, nt = METHOD == NL_DIV ? 192 : 384;
    const dim3 grid((unsigned)((g.M + YW_MODES - 1) / YW_MODES), n);
    if (nrt <= 1) k_yinv_w<METHOD, 1><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else if (nrt <= 2) k_yinv_w<METHOD, 2><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else if (nrt <= 3) k_yinv_w<METHOD, 3><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else if (nrt <= 4) k_yinv_w<METHOD, 4><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else if (nrt <= 6) k_yinv_w<METHOD, 6><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else if (nrt <= 8) k_yinv_w<METHOD, 8><<<grid, nt, 0, st>>>(g, ym, tab, yp, sy);
    else {
        const dim3 tg((unsigned)(((g.M + YB_M - 1) / YB_M) * ((g.Ny + YB_R - 1) / YB_R)), n);
        k_yinv_t<METHOD><<<tg, 256, 0, st>>>(g, ym, tab, yp, sy);
    }
}

// xfold: the skew-symmetric fields come folded from k_xwfwd (f_c with its
// x-derivative term in fields 0-2, u_cu_1 in 3 + c)
template <int METHOD>
static void launch_yfwd(const Geom& g, const YMats& ym, int n, int have_g, const double2* fo, long sy,
                        const NLMember* tab, hipStream_t st, bool xfold = false) {
    const int nrt = (g.Ny / 2 + 15) / 16, xf = xfold ? 1 : 0;
    const dim3 grid((unsigned)((g.M + YW_MODES - 1) / YW_MODES), n);
    if (nrt <= 1) k_yfwd_w<METHOD, 1><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else if (nrt <= 2) k_yfwd_w<METHOD, 2><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else if (nrt <= 3) k_yfwd_w<METHOD, 3><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else if (nrt <= 4) k_yfwd_w<METHOD, 4><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else if (nrt <= 6) k_yfwd_w<METHOD, 6><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else if (nrt <= 8) k_yfwd_w<METHOD, 8><<<grid, 192, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    else {
        const dim3 tg((unsigned)(((g.M + YB_M - 1) / YB_M) * ((g.Ny + YB_R - 1) / YB_R)), n);
        k_yfwd_t<METHOD><<<tg, 256, 0, st>>>(g, ym, have_g, fo, sy, tab, xf);
    }
}

// LDS budget per block (set from the device's opt-in limit at start-up)
static size_t g_lds_budget = 64 * 1024;
void set_lds_budget(size_t bytes) { g_lds_budget = bytes; }

// real physical lines per (y, x) line the z-pass keeps in LDS
static constexpr int zpass_nphys(int method) { return method == NL_ROT ? 6 : (method == NL_DIV ? 5 : 7); }

// largest z-pass lines per block that fit the LDS budget (power of two <= 32)
static int zpass_xb_max(const Geom& g, int nphys) {
    const long per = 32L * lpitch(g.Nz) + 8L * g.Nz * nphys;  // two complex work lines + nphys real lines
    const long avail = (long)g_lds_budget - 16L * g.Nz;
    int xb = 1;
    while (xb < 32 && xb * 2 <= g.Nx && (long)(xb * 2) * per <= avail) xb *= 2;
    return xb;
}

// largest x-FFT lines per block that fit (power of two <= 16, at most the kz lines)
static int xfft_lines_max(const Geom& g) {
    const long per = 32L * lpitch(g.Nx);  // two complex work lines
    const long avail = (long)g_lds_budget - 16L * g.Nx;
    int xl = 1;
    while (xl < 16 && xl * 2 <= g.nkz && (long)(xl * 2) * per <= avail) xl *= 2;
    return xl;
}

static int zpass_xb(const Geom& g, int nphys, const NLTune& tn) {
    const int m = zpass_xb_max(g, nphys);
    return tn.zxb > 0 ? std::min(tn.zxb, m) : m;
}
static int xfft_lines(const Geom& g, const NLTune& tn) {
    const int m = xfft_lines_max(g);
    return tn.xl > 0 ? std::min(tn.xl, m) : std::min(4, m);
}

#define CFG_FFT_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// fold (forward only): the skew-symmetric layout of k_xwfwd (8 z-pass fields in,
// 6 out, the x-derivative part of the divergence added); only the wavefront
// FFT lengths have it. Returns whether it folded.
template <int NX>
static bool launch_x(bool inverse, const Geom& g, const FFTPlan& px, int nsrc, int nf, int n, const double2* in,
                     long sin, double2* out, long sout, const NLTune& tn, hipStream_t st, bool fold) {
    if constexpr (zwave_size(NX)) {
        using S = ZW<NX>;
        const int wb = (!inverse && tn.xwbf > 0) ? tn.xwbf : (tn.xwb > 0 ? tn.xwb : 4);
        const long waves = (long)g.Ny * ((g.nkz + S::LPW - 1) / S::LPW);
        const size_t lds = (size_t)(NX + wb * S::LPW * S::LBUF) * sizeof(double2);
        if (inverse) {
            const dim3 grid((unsigned)((waves + wb - 1) / wb), nf, n);
            allow_lds(k_xwinv<NX>, lds);
            k_xwinv<NX><<<grid, 64 * wb, lds, st>>>(g, px.tw, nsrc, in, sin, out, sout);
        } else {
            if (fold && nf != 8) throw std::runtime_error("x-FFT fold: the skew-symmetric layout has 8 fields");
            const dim3 grid((unsigned)((waves + wb - 1) / wb), fold ? 5 : nf, n);
            allow_lds(k_xwfwd<NX>, lds);
            const int xcd = (g.nkz * (int)sizeof(double2)) % 128 != 0;
            k_xwfwd<NX><<<grid, 64 * wb, lds, st>>>(g, px.tw, in, sin, out, sout, fold ? 1 : 0, xcd);
        }
        return fold && !inverse;
    }
    const int XL = xfft_lines(g, tn);
    const int nchunk = (g.nkz + XL - 1) / XL;
    const size_t lds = (size_t)(g.Nx + 2 * XL * lpitch(g.Nx)) * sizeof(double2);
    const dim3 grid(g.Ny * nchunk, nf, n);
    if (inverse) {
        allow_lds(k_xinv<NX>, lds);
        k_xinv<NX><<<grid, 256, lds, st>>>(g, px, XL, nsrc, in, sin, out, sout);
    } else {
        allow_lds(k_xfwd<NX>, lds);
        k_xfwd<NX><<<grid, 256, lds, st>>>(g, px, XL, in, sin, out, sout);
    }
    return false;
}

// in/out strides are per member, in double2. Returns whether the forward
// transform folded the skew-symmetric layout (launch_x).
static bool launch_xfft(bool inverse, const Geom& g, const FFTPlan& px, int nsrc, int nf, int n, const double2* in,
                        long sin, double2* out, long sout, const NLTune& tn, hipStream_t st, bool fold = false) {
    switch (g.Nx) {
#define CASE_X(N) \
    case N: return launch_x<N>(inverse, g, px, nsrc, nf, n, in, sin, out, sout, tn, st, fold);
        CFG_FFT_SIZES(CASE_X)
#undef CASE_X
        default: return launch_x<0>(inverse, g, px, nsrc, nf, n, in, sin, out, sout, tn, st, fold);
    }
}

template <int METHOD, int NZ>
static void launch_z_t(const Geom& g, const FFTPlan& pz, int n, const NLMember* tab, int want_g, const double2* xp,
                       double2* zo, long sx, const NLTune& tn, hipStream_t st) {
    if constexpr (zwave_size(NZ)) {
        using S = ZW<NZ>;
        const long waves = (long)g.Ny * ((g.Nx + S::LPW - 1) / S::LPW);
        if (tn.zstage < 0 ? zwave_staged(METHOD, NZ) : tn.zstage != 0) {
            if (g.nkz > NZ / 2) throw std::runtime_error("z pass: more live kz bins than Nz/2");
            const dim3 grid((unsigned)((waves + ZWAVE_WB - 1) / ZWAVE_WB), n);
            k_zwave<METHOD, NZ, true><<<grid, 64 * ZWAVE_WB, 0, st>>>(g, pz.tw, tab, want_g, xp, zo, sx);
            return;
        }
        const int wb = tn.zwb > 0 ? tn.zwb : 4;
        const size_t lds = (size_t)(NZ + wb * S::LPW * S::LBUF) * sizeof(double2);
        const dim3 grid((unsigned)((waves + wb - 1) / wb), n);
        allow_lds(k_zwave<METHOD, NZ, false>, lds);
        k_zwave<METHOD, NZ, false><<<grid, 64 * wb, lds, st>>>(g, pz.tw, tab, want_g, xp, zo, sx);
        return;
    }
    const int nphys = zpass_nphys(METHOD);
    const int XB = zpass_xb(g, nphys, tn);
    const size_t lds = (size_t)g.Nz * sizeof(double2) +
                       (size_t)XB * (2 * lpitch(g.Nz) * sizeof(double2) + (size_t)g.Nz * nphys * sizeof(double));
    const dim3 grid(g.Ny * ((g.Nx + XB - 1) / XB), n);
    allow_lds(k_zpass<METHOD, NZ>, lds);
    k_zpass<METHOD, NZ><<<grid, 256, lds, st>>>(g, pz, XB, tab, want_g, xp, zo, sx);
}

template <int METHOD>
static void launch_z(const Geom& g, const FFTPlan& pz, int n, const NLMember* tab, int want_g, const double2* xp,
                     double2* zo, long sx, const NLTune& tn, hipStream_t st) {
    switch (g.Nz) {
#define CASE_Z(N) \
    case N: launch_z_t<METHOD, N>(g, pz, n, tab, want_g, xp, zo, sx, tn, st); return;
        CFG_FFT_SIZES(CASE_Z)
#undef CASE_Z
        default: launch_z_t<METHOD, 0>(g, pz, n, tab, want_g, xp, zo, sx, tn, st);
    }
}

static void launch_z_any(int method, const Geom& g, const FFTPlan& pz, int n, const NLMember* tab, int want_g,
                         const double2* xp, double2* zo, long sx, const NLTune& tn, hipStream_t st) {
    switch (method) {
        case NL_ROT: launch_z<NL_ROT>(g, pz, n, tab, want_g, xp, zo, sx, tn, st); break;
        case NL_CONV: launch_z<NL_CONV>(g, pz, n, tab, want_g, xp, zo, sx, tn, st); break;
        case NL_SKEW: launch_z<NL_SKEW>(g, pz, n, tab, want_g, xp, zo, sx, tn, st); break;
        default: launch_z<NL_DIV>(g, pz, n, tab, want_g, xp, zo, sx, tn, st);
    }
}

template <int NZ>
static void launch_zcfl_t(const Geom& g, const FFTPlan& pz, int n, const NLMember* tab, const double2* xp, long sx,
                          double* part, int XB, hipStream_t st) {
    const size_t lds = (size_t)g.Nz * sizeof(double2) +
                       (size_t)XB * (2 * lpitch(g.Nz) * sizeof(double2) + (size_t)g.Nz * 3 * sizeof(double));
    const dim3 grid(g.Ny * ((g.Nx + XB - 1) / XB), n);
    allow_lds(k_zcfl<NZ>, lds);
    k_zcfl<NZ><<<grid, 256, lds, st>>>(g, pz, XB, tab, xp, sx, part);
}

void launch_cfl(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int n, const NLMember* tab,
                const NLWork& w, double* out, hipStream_t st) {
    const long sy = nlwork_y_doubles(g) / 2, sx = nlwork_x_doubles(g) / 2;
    double2* yp = reinterpret_cast<double2*>(w.yp);
    double2* xp = reinterpret_cast<double2*>(w.xp);
    double* part = w.zo;  // n x blocks partial maxima
    launch_yinv<NL_DIV>(g, ym, n, tab, yp, sy, st);
    launch_xfft(true, g, px, 3, 3, n, yp, sy, xp, sx, NLTune{}, st);
    const int XB = zpass_xb_max(g, 3);
    const int nb = g.Ny * ((g.Nx + XB - 1) / XB);
    switch (g.Nz) {
#define CASE_Z(N) \
    case N: launch_zcfl_t<N>(g, pz, n, tab, xp, sx, part, XB, st); break;
        CFG_FFT_SIZES(CASE_Z)
#undef CASE_Z
        default: launch_zcfl_t<0>(g, pz, n, tab, xp, sx, part, XB, st);
    }
    k_cfl_reduce<<<n, 256, 0, st>>>(nb, part, out);
}

// Nonlinear term of n members of one grid class. rotating: some member has a
// nonzero rotation (only then does the divergence form compute g).
void launch_nonlinear(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int method, int n,
                      const NLMember* tab, bool rotating, const NLWork& w, const NLTune& tn, hipStream_t st,
                      KTimer* kt) {
    KTimer dummy;
    KTimer& T = kt ? *kt : dummy;
    const long sy = nlwork_y_doubles(g) / 2, sx = nlwork_x_doubles(g) / 2;  // per-member strides in double2
    double2* yp = reinterpret_cast<double2*>(w.yp);
    double2* xp = reinterpret_cast<double2*>(w.xp);
    double2* zo = reinterpret_cast<double2*>(w.zo);
    double2* fo = reinterpret_cast<double2*>(w.fo);
    int nsrc = 3, nxout = 3;
    hipEvent_t ev = T.begin();
    switch (method) {
        case NL_ROT:
            launch_yinv<NL_ROT>(g, ym, n, tab, yp, sy, st);
            nsrc = 6; nxout = 6;
            break;
        case NL_CONV:
            launch_yinv<NL_CONV>(g, ym, n, tab, yp, sy, st);
            nsrc = 6; nxout = 9;
            break;
        case NL_SKEW:
            launch_yinv<NL_SKEW>(g, ym, n, tab, yp, sy, st);
            nsrc = 6; nxout = 9;
            break;
        default:
            launch_yinv<NL_DIV>(g, ym, n, tab, yp, sy, st);
            nsrc = 3; nxout = 3;
    }
    T.end(K_YINV, ev);
    ev = T.begin();
    // d/dx outputs (fields nsrc..) derive from fields 0..2 only
    launch_xfft(true, g, px, nsrc == 6 && nxout == 9 ? 6 : nxout, nxout, n, yp, sy, xp, sx, tn, st);
    T.end(K_XINV, ev);
    const int want_g = (method != NL_DIV) || rotating;
    ev = T.begin();
    launch_z_any(method, g, pz, n, tab, want_g, xp, zo, sx, tn, st);
    T.end(K_ZPASS, ev);
    ev = T.begin();
    // forward x-FFT of the z-pass outputs that exist: fields 0-2 (g) when
    // want_g, fields 3-8 (u_i u_j) in the divergence form, 3-7 in the
    // skew-symmetric one; one launch when both exist (the fields are
    // contiguous)
    const bool uu = method == NL_DIV || method == NL_SKEW;
    bool xfold = false;  // the skew-symmetric x-derivative terms already in fields 0-2 (k_xwfwd)
    if (want_g)
        xfold = launch_xfft(false, g, px, 0, method == NL_SKEW ? 8 : (uu ? 9 : 3), n, zo, sx, fo, sy, tn, st,
                            method == NL_SKEW);
    else if (uu) {
        const long off = 3L * g.Ny * g.Nx * g.nkz;
        const long offo = 3L * g.Ny * g.M;
        launch_xfft(false, g, px, 0, 6, n, zo + off, sx, fo + offo, sy, tn, st);
    }
    T.end(K_XFWD, ev);
    ev = T.begin();
    switch (method) {
        case NL_ROT: launch_yfwd<NL_ROT>(g, ym, n, 1, fo, sy, tab, st); break;
        case NL_CONV: launch_yfwd<NL_CONV>(g, ym, n, 1, fo, sy, tab, st); break;
        case NL_SKEW: launch_yfwd<NL_SKEW>(g, ym, n, 1, fo, sy, tab, st, xfold); break;
        default: launch_yfwd<NL_DIV>(g, ym, n, want_g, fo, sy, tab, st);
    }
    T.end(K_YFWD, ev);
}

// Autotuner of the nonlinear-term tilings: each candidate block shape of the
// z-pass and of the x-FFTs (wavefronts per block for the wavefront FFTs,
// lines per block for the LDS ones) is timed (HIP events, best of 3 after a
// warm launch) on the caller's work buffers and member table. The tiling only
// changes which threads and blocks do a line's butterflies, never the
// arithmetic, so every choice gives bit-identical results.
NLTune tune_nonlinear(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int method, int n,
                      const NLMember* tab, bool rotating, const NLWork& w, hipStream_t st) {
    (void)ym;
    const long sy = nlwork_y_doubles(g) / 2, sx = nlwork_x_doubles(g) / 2;
    double2* yp = reinterpret_cast<double2*>(w.yp);
    double2* xp = reinterpret_cast<double2*>(w.xp);
    double2* zo = reinterpret_cast<double2*>(w.zo);
    const int want_g = (method != NL_DIV) || rotating;
    const int nphys = zpass_nphys(method);
    const int nxout = (method == NL_ROT) ? 6 : (method == NL_DIV ? 3 : 9);
    const int nsrc = (method == NL_DIV) ? 3 : 6;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return NLTune{};
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return NLTune{};
    }
    // Every candidate runs once to warm up, then all are timed in 5 interleaved
    // rounds (so clock ramps and neighbours' load fall on all of them alike);
    // each keeps its best time.
    auto pick = [&](const std::vector<NLTune>& cands, auto&& launch, const char* what) {
        for (const NLTune& c : cands) launch(c);
        std::vector<float> best(cands.size(), 1e30f);
        for (int r = 0; r < 5; ++r)
            for (size_t i = 0; i < cands.size(); ++i) {
                (void)hipEventRecord(a, st);
                launch(cands[i]);
                (void)hipEventRecord(b, st);
                (void)hipEventSynchronize(b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                best[i] = std::min(best[i], ms);
            }
        size_t k = 0;
        for (size_t i = 1; i < cands.size(); ++i)
            if (best[i] < best[k]) k = i;
        if (std::getenv("CFGPU_TUNE_LOG")) {
            std::fprintf(stderr, "cfg tune %s %dx%dx%d form %d members %d:", what, g.Nx, g.Ny, g.Nz, method, n);
            for (size_t i = 0; i < cands.size(); ++i)
                std::fprintf(stderr, " [zwb %d zstage %d zxb %d xwb %d xwbf %d xl %d] %.1f us%s", cands[i].zwb,
                             cands[i].zstage, cands[i].zxb, cands[i].xwb, cands[i].xwbf, cands[i].xl, best[i] * 1e3f,
                             i == k ? " *" : "");
            std::fprintf(stderr, "\n");
        }
        return cands[k];
    };
    NLTune out;
    std::vector<NLTune> zc;
    if (zwave_size(g.Nz)) {  // wavefront z pass: the staged form, or wavefronts per block
        // the grid's and form's choice (zwave_staged), or the one CFGPU_ZWAVE_STAGED=1 / 0 names
        const char* force = std::getenv("CFGPU_ZWAVE_STAGED");
        const bool staged = force ? force[0] != '0' : zwave_staged(method, g.Nz), unstaged = !staged;
        NLTune c;
        if (staged) {
            c.zwb = ZWAVE_WB;
            c.zstage = 1;
            zc.push_back(c);
        }
        for (int wb = 1; unstaged && wb <= 4; wb *= 2) {
            c.zwb = wb;
            c.zstage = 0;
            zc.push_back(c);
        }
    } else {  // LDS z pass: lines per block
        for (int xb = 1; xb <= zpass_xb_max(g, nphys); xb *= 2) {
            NLTune c;
            c.zxb = xb;
            zc.push_back(c);
        }
    }
    if (!zc.empty()) {
        const NLTune z = pick(zc, [&](const NLTune& c) { launch_z_any(method, g, pz, n, tab, want_g, xp, zo, sx, c, st); },
                              "z");
        out.zwb = z.zwb;
        out.zstage = z.zstage;
        out.zxb = z.zxb;
    }
    std::vector<NLTune> xc;  // x-FFTs, timed on the inverse
    if (zwave_size(g.Nx)) {  // wavefront x-FFTs: wavefronts per block
        for (int wb = 1; wb <= 4; wb *= 2) {
            NLTune c;
            c.xwb = wb;
            xc.push_back(c);
        }
    } else {  // LDS x-FFTs: lines per block
        for (int xl = 1; xl <= xfft_lines_max(g); xl *= 2) {
            NLTune c;
            c.xl = xl;
            xc.push_back(c);
        }
    }
    if (!xc.empty()) {
        const NLTune x = pick(xc, [&](const NLTune& c) {
            launch_xfft(true, g, px, nsrc == 6 && nxout == 9 ? 6 : nxout, nxout, n, yp, sy, xp, sx, c, st);
        }, "x");
        out.xwb = x.xwb;
        out.xl = x.xl;
    }
    if (zwave_size(g.Nx) && want_g) {  // the forward wavefront x-FFT on its own (it folds the skew form)
        const int nzf = method == NL_SKEW ? 8 : (method == NL_DIV ? 9 : 3);  // as launch_nonlinear
        std::vector<NLTune> fc;
        for (int wb = 1; wb <= 4; wb *= 2) {
            NLTune c;
            c.xwbf = wb;
            fc.push_back(c);
        }
        const NLTune f = pick(fc, [&](const NLTune& c) {
            launch_xfft(false, g, px, 0, nzf, n, zo, sx, yp, sy, c, st, method == NL_SKEW);
        }, "xfwd");
        out.xwbf = f.xwbf;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return out;
}

}  // namespace cfg
