// Host-side infrastructure shared by the single-field and batched DNS
// runtimes: errors, device buffers, Chebyshev helpers, the laminar base
// profile, per-grid constant tables (FFT twiddles, y operators) and the
// compact <-> channelflow layout conversion.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_set>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/cfgpu.h"
#include "cfg_kernels.hpp"
#include "cfg_math.hpp"

namespace cfg {

inline std::string& last_error() {
    static thread_local std::string s;
    return s;
}

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) throw Error(CFG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DBuf {
    double* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        std::swap(p, o.p);
        std::swap(n, o.n);
        return *this;
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) HIPCHK(hipMalloc(&p, count * sizeof(double)));
    }
    void zero(hipStream_t st) {
        if (n) HIPCHK(hipMemsetAsync(p, 0, n * sizeof(double), st));
    }
};

// ---------------------------------------------------------------------------
// host-side Chebyshev helpers (chebyshev.cpp:262-303, 746-754)
// ---------------------------------------------------------------------------
inline long double cos_pi_frac(long num, long den) {  // cos(pi num/den), reduced exactly
    long r = num % (2 * den);
    if (r < 0) r += 2 * den;
    return cosl(3.141592653589793238462643383279502884L * (long double)r / (long double)den);
}

inline std::vector<double> cheb_points(int N, double a, double b) {
    std::vector<double> x(N);
    const double piN = 3.14159265358979323846 / (N - 1);
    const double radius = (b - a) / 2, center = (b + a) / 2;
    for (int j = 0; j < N; ++j) x[j] = center + radius * cos(j * piN);
    return x;
}

// values at the Chebyshev points -> coefficients (the normalised DCT-I)
inline std::vector<double> cheb_coeffs(const std::vector<double>& v) {
    const int N = (int)v.size();
    std::vector<double> c(N);
    for (int k = 0; k < N; ++k) {
        long double s = 0.0L;
        for (int n = 0; n < N; ++n) {
            const long double e = (n == 0 || n == N - 1) ? 1.0L : 2.0L;
            s += e * (long double)v[n] * cos_pi_frac((long)n * k, N - 1);
        }
        const long double sk = (k == 0 || k == N - 1) ? 0.5L : 1.0L;
        c[k] = (double)(sk * s / (long double)(N - 1));
    }
    return c;
}

inline std::vector<double> host_diff(const std::vector<double>& u, double L) {
    std::vector<double> du(u.size());
    std::vector<double> uu(u);
    cheb_diff(SV{uu.data(), 1}, SV{du.data(), 1}, (int)u.size(), L);
    return du;
}

inline double host_eval_a(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_eval_a(SV{uu.data(), 1}, (int)u.size());
}
inline double host_eval_b(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_eval_b(SV{uu.data(), 1}, (int)u.size());
}
inline double host_mean(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_mean(SV{uu.data(), 1}, (int)u.size());
}

// laminar profile (nse.cpp:823-943)
inline std::vector<double> laminar_profile(double nu, int constraint, double dPdx, double Ubulk, double Vsuck,
                                           double a, double b, double ua, double ub, int Ny) {
    std::vector<double> u(Ny, 0.0);
    const double H = b - a;
    const double mu = Vsuck * H / nu;
    auto sq = [](double x) { return x * x; };
    if (std::abs(mu) == 0.0) {
        if (constraint == CFG_BulkVelocity) {
            u[0] = 0.125 * (ub + ua) + 0.75 * Ubulk;
            u[1] = 0.5 * (ub - ua);
            u[2] = 0.375 * (ub + ua) - 0.75 * Ubulk;
        } else {
            dPdx *= sq((b - a) / 2);
            u[0] = 0.5 * (ub + ua) - 0.25 * dPdx / nu;
            u[1] = 0.5 * (ub - ua);
            u[2] = 0.25 * dPdx / nu;
        }
        return u;
    }
    const double ub_ua = ub - ua;
    const double expm1_mu = std::expm1(-mu);
    double dPdx_HH_nu = dPdx * sq(H) / nu;
    const std::vector<double> y = cheb_points(Ny, a, b);
    if (std::abs(mu) > 1e-01) {
        if (constraint == CFG_BulkVelocity) {
            const double k = -1.0 / expm1_mu - 1 / mu;
            dPdx_HH_nu = mu * (Ubulk - ua - ub_ua * k) / (0.5 - k);
        }
        for (int i = 0; i < Ny; i++) {
            const double yhat = (y[i] - a) / H;
            const double e = std::expm1(-mu * yhat);
            u[i] = ua + (ub_ua)*e / expm1_mu + dPdx_HH_nu * (yhat - e / expm1_mu) / mu;
        }
    } else {
        if (constraint == CFG_BulkVelocity) {
            const double mu2 = mu * mu, A = ub - ua, B = ua + ub - 2 * Ubulk;
            dPdx_HH_nu = 6 * B + mu * (A + mu * (B / 10 + mu2 * (-B / 1400 + mu2 * B / 126000)));
        }
        const double o2 = 1.0 / 2, o12 = 1.0 / 12, o24 = 1.0 / 24, o720 = 1.0 / 720, o1440 = 1.0 / 1440,
                     o30240 = 1.0 / 30240, o120960 = 1.0 / 120960;
        for (int i = 0; i < Ny; i++) {
            const double yh = (y[i] - a) / H, yh2 = yh * yh, yh3 = yh2 * yh, yh4 = yh3 * yh, yh_1 = yh - 1;
            const double t1 =
                yh + mu * (-o2 * yh * yh_1 +
                           mu * (o12 * yh * yh_1 * (2 * yh - 1) +
                                 mu * (-o24 * sq(yh * yh_1) +
                                       mu * (o720 * yh * yh_1 * (2 * yh - 1) * (3 * yh2 - 3 * yh - 1) +
                                             mu * (-o1440 * sq(yh * yh_1) * (2 * yh2 - 2 * yh - 1) +
                                                   mu * (o30240 * yh * yh_1 * (2 * yh - 1) *
                                                         (3 * yh4 - 6 * yh3 + 3 * yh + 1)))))));
            const double t2 =
                o2 * yh * yh_1 + mu * (-o12 * yh * yh_1 * (2 * yh - 1) +
                                       mu * (o24 * sq(yh * yh_1) +
                                             mu * (-o720 * yh * yh_1 * (2 * yh - 1) * (3 * yh2 - 3 * yh - 1) +
                                                   mu * (o1440 * sq(yh * yh_1) * (2 * yh2 - 2 * yh - 1) +
                                                         mu * (-o30240 * yh * yh_1 * (2 * yh - 1) *
                                                                   (3 * yh4 - 6 * yh3 + 3 * yh + 1) +
                                                               mu * (o120960 * sq(yh * yh_1) *
                                                                     (3 * yh4 - 6 * yh3 + 4 * yh + 2)))))));
            u[i] = ua + ub_ua * t1 + dPdx_HH_nu * t2;
        }
    }
    return cheb_coeffs(u);
}

// ---------------------------------------------------------------------------
// per-geometry constant tables: FFT plans, y operators
// ---------------------------------------------------------------------------
// FFT plan of length N (radices 4, 2, 3) and its twiddles exp(-2 pi i j/N),
// evaluated in long double; the plan's tw pointer is set by the caller.
inline FFTPlan fft_plan(int N, std::vector<double>& tw) {
    FFTPlan plan{};
    int n = N, np = 0;
    while (n % 4 == 0) { plan.radix[np++] = 4; n /= 4; }
    while (n % 2 == 0) { plan.radix[np++] = 2; n /= 2; }
    while (n % 3 == 0) { plan.radix[np++] = 3; n /= 3; }
    if (n != 1) throw Error(CFG_ERR_ARG, "grid size " + std::to_string(N) + " is not of the form 2^a 3^b");
    plan.N = N;
    plan.npass = np;
    tw.assign(2 * N, 0.0);
    for (int j = 0; j < N; ++j) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * j / N;
        tw[2 * j] = (double)cosl(ang);
        tw[2 * j + 1] = (double)sinl(ang);
    }
    return plan;
}

inline Geom make_geom(const cfg_grid& gr, bool dealias) {
    Geom g{};
    g.Nx = gr.Nx;
    g.Ny = gr.Ny;
    g.Nz = gr.Nz;
    g.kxd = dealias ? gr.Nx / 3 - 1 : gr.Nx / 2 - 1;
    g.kzd = dealias ? gr.Nz / 3 - 1 : gr.Nz / 2 - 1;
    g.nkx = 2 * g.kxd + 1;
    g.nkz = g.kzd + 1;
    g.M = g.nkx * g.nkz;
    g.Lx = gr.Lx;
    g.Lz = gr.Lz;
    g.a = gr.a;
    g.b = gr.b;
    return g;
}

// y operators of a grid, row-major Ny x Ny (see YMats), summed in long double
struct YTables {
    std::vector<double> T, TD, Ti, DTi;
    std::vector<long double> tcheb;  // T_k(y_n) = cos(pi n k / (Ny-1)), [n][k]
};

inline YTables y_tables(int N, double a, double b) {
    YTables yt;
    const long double L = (long double)(b - a);
    std::vector<long double> t(N * N), d(N * N, 0.0L), ti(N * N);
    for (int y = 0; y < N; ++y)
        for (int k = 0; k < N; ++k) t[y * N + k] = cos_pi_frac((long)y * k, N - 1);
    for (int k = 0; k < N; ++k)
        for (int y = 0; y < N; ++y) {
            const long double sk = (k == 0 || k == N - 1) ? 0.5L : 1.0L;
            const long double ey = (y == 0 || y == N - 1) ? 1.0L : 2.0L;
            ti[k * N + y] = sk * ey * t[y * N + k] / (long double)(N - 1);
        }
    // d/dy on coefficients: du[n] = sum_{m>n, m-n odd} (4/L) m u[m], du[0] halved
    for (int n = 0; n < N; ++n)
        for (int m = n + 1; m < N; m += 2) d[n * N + m] = (4.0L / L) * m * (n == 0 ? 0.5L : 1.0L);
    yt.T.resize(N * N);
    yt.TD.resize(N * N);
    yt.Ti.resize(N * N);
    yt.DTi.resize(N * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            long double sd = 0.0L, sdi = 0.0L;
            for (int k = 0; k < N; ++k) {
                sd += t[i * N + k] * d[k * N + j];
                sdi += d[i * N + k] * ti[k * N + j];
            }
            yt.T[i * N + j] = (double)t[i * N + j];
            yt.TD[i * N + j] = (double)sd;
            yt.Ti[i * N + j] = (double)ti[i * N + j];
            yt.DTi[i * N + j] = (double)sdi;
        }
    yt.tcheb = std::move(t);
    return yt;
}

// y values of a coefficient profile and of its derivative (long double sums)
inline std::vector<double> eval_profile(const Geom& g, const std::vector<long double>& tcheb,
                                        const std::vector<double>& c, bool deriv) {
    const int N = g.Ny;
    std::vector<double> cd = deriv ? host_diff(c, g.b - g.a) : c;
    std::vector<double> v(N);
    for (int y = 0; y < N; ++y) {
        long double s = 0.0L;
        for (int k = 0; k < N; ++k) s += tcheb[(size_t)y * N + k] * (long double)cd[k];
        v[y] = (double)s;
    }
    return v;
}

struct FFTTable {
    FFTPlan plan{};
    DBuf tw;
    void build(int N) {
        std::vector<double> h;
        plan = fft_plan(N, h);
        tw.alloc(2 * N);
        HIPCHK(hipMemcpy(tw.p, h.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice));
        plan.tw = reinterpret_cast<const double2*>(tw.p);
    }
};

struct Geometry {
    Geom g{};
    int dealias = 0;
    FFTTable fx, fz;
    DBuf T, TD, Ti, DTi;
    DBuf HB;  // rows of the Chebyshev-tau operator B of the Helmholtz systems, [3][Ny]
    std::vector<long double> tcheb;  // T_k(y_n) = cos(pi n k / (Ny-1)), [n][k], for host profile evaluation

    YMats ymats() const { return YMats{T.p, TD.p, Ti.p, DTi.p}; }

    void build(const cfg_grid& gr, bool dealias_xz) {
        dealias = dealias_xz ? 1 : 0;
        g = make_geom(gr, dealias_xz);
        fx.build(g.Nx);
        fz.build(g.Nz);
        const int N = g.Ny;
        YTables yt = y_tables(N, gr.a, gr.b);
        tcheb = yt.tcheb;
        auto up = [&](DBuf& b, const std::vector<double>& h) {
            b.alloc(h.size());
            HIPCHK(hipMemcpy(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        };
        up(T, yt.T);
        up(TD, yt.TD);
        up(Ti, yt.Ti);
        up(DTi, yt.DTi);
        // B of nu u'' - lambda u = f on row n >= 2 (helmholtz.cpp:79-95 with the
        // c_n and beta_n of helmholtz.h:65-66): c_{n-2}/(4n(n-1)) f[n-2]
        // - beta_n/(2(n^2-1)) f[n] + beta_{n+2}/(4n(n+1)) f[n+2], the last
        // term only where row n + 2 belongs to the same parity system
        std::vector<double> hb(3 * N, 0.0);
        const int Nc = N - 1;
        auto cn = [&](int n) { return (n == 0 || n == Nc) ? 2 : 1; };
        auto beta = [&](int n) { return (n > Nc - 2) ? 0 : 1; };
        for (int n = 2; n < N; ++n) {
            hb[n] = ((double)cn(n - 2)) / (4 * n * (n - 1));
            hb[N + n] = -((double)beta(n)) / (2 * (n * n - 1));
            hb[2 * N + n] = (n + 2 < N && beta(n + 2)) ? 1.0 / (4 * n * (n + 1)) : 0.0;
        }
        up(HB, hb);
    }

    std::vector<double> eval_profile(const std::vector<double>& c, bool deriv) const {
        return cfg::eval_profile(g, tcheb, c, deriv);
    }
};

// compact <-> channelflow layout
inline void pack(const Geom& g, int nd, const double* ref, std::vector<double>& out) {
    const int Mz = g.Nz / 2 + 1;
    out.assign((size_t)2 * nd * g.Ny * g.M, 0.0);
    if (!ref) return;
    for (int c = 0; c < nd; ++c)
        for (int ny = 0; ny < g.Ny; ++ny)
            for (int ikz = 0; ikz < g.nkz; ++ikz)
                for (int ikx = 0; ikx < g.nkx; ++ikx) {
                    const int kx = ikx <= g.kxd ? ikx : ikx - g.nkx;
                    const int mx = kx >= 0 ? kx : kx + g.Nx;
                    const size_t src = (size_t)ikz + Mz * ((size_t)mx + g.Nx * ((size_t)ny + g.Ny * c));
                    const size_t dst = (((size_t)c * g.Ny + ny) * g.nkz + ikz) * g.nkx + ikx;
                    out[2 * dst] = ref[2 * src];
                    out[2 * dst + 1] = ref[2 * src + 1];
                }
}

inline void unpack(const Geom& g, int nd, const std::vector<double>& in, double* ref) {
    const int Mz = g.Nz / 2 + 1;
    std::memset(ref, 0, sizeof(double) * 2 * (size_t)nd * g.Ny * g.Nx * Mz);
    for (int c = 0; c < nd; ++c)
        for (int ny = 0; ny < g.Ny; ++ny)
            for (int ikz = 0; ikz < g.nkz; ++ikz)
                for (int ikx = 0; ikx < g.nkx; ++ikx) {
                    const int kx = ikx <= g.kxd ? ikx : ikx - g.nkx;
                    const int mx = kx >= 0 ? kx : kx + g.Nx;
                    const size_t dst = (size_t)ikz + Mz * ((size_t)mx + g.Nx * ((size_t)ny + g.Ny * c));
                    const size_t src = (((size_t)c * g.Ny + ny) * g.nkz + ikz) * g.nkx + ikx;
                    ref[2 * dst] = in[2 * src];
                    ref[2 * dst + 1] = in[2 * src + 1];
                }
}

inline int smooth235(int n) {
    for (int p : {2, 3})
        while (n % p == 0) n /= p;
    return n == 1;
}

inline void check_grid(const cfg_grid& gr) {
    if (gr.Nx < 4 || gr.Nz < 4 || gr.Ny < 5 || gr.Ny % 2 == 0)
        throw Error(CFG_ERR_ARG, "grid must have Nx, Nz >= 4 and odd Ny >= 5 (Chebyshev-tau needs odd Ny)");
    if (gr.Nx % 2 || gr.Nz % 2) throw Error(CFG_ERR_ARG, "Nx and Nz must be even");
    if (!smooth235(gr.Nx) || !smooth235(gr.Nz)) throw Error(CFG_ERR_ARG, "Nx and Nz must be of the form 2^a 3^b");
    if (gr.Ny > 383) throw Error(CFG_ERR_ARG, "Ny > 383 is not supported by the tau-solve kernels");
    if (gr.Nx > 256 || gr.Nz > 512) throw Error(CFG_ERR_ARG, "Nx <= 256 and Nz <= 512 are supported");
    if (!(gr.b > gr.a) || !(gr.Lx > 0) || !(gr.Lz > 0)) throw Error(CFG_ERR_ARG, "invalid box");
}

inline void check_flags(const cfg_flags& f) {
    if (f.dealiasing == CFG_DealiasY || f.dealiasing == CFG_DealiasXYZ)
        throw Error(CFG_ERR_ARG, "DealiasY / DealiasXYZ (Chebyshev-direction dealiasing) is not supported");
    if (f.dealiasing != CFG_NoDealiasing && f.dealiasing != CFG_DealiasXZ) throw Error(CFG_ERR_ARG, "bad dealiasing");
    if (f.nonlinearity < 0 || f.nonlinearity > CFG_Alternating_) throw Error(CFG_ERR_ARG, "unsupported nonlinearity");
    if (f.rotation != 0.0 && f.nonlinearity != CFG_Rotational && f.nonlinearity != CFG_SkewSymmetric)
        throw Error(CFG_ERR_ARG, "rotation is supported with Rotational and SkewSymmetric nonlinearity only");
    if (f.timestepping < CFG_CNFE1 || f.timestepping > CFG_SBDF4) throw Error(CFG_ERR_ARG, "bad timestepping");
    if (f.initstepping < CFG_CNFE1 || f.initstepping > CFG_SBDF4) throw Error(CFG_ERR_ARG, "bad initstepping");
    if (!(f.nu > 0) || !(f.dt > 0)) throw Error(CFG_ERR_ARG, "nu and dt must be positive");
}


// ---------------------------------------------------------------------------
// C-boundary error handling: no exception crosses the ABI
// ---------------------------------------------------------------------------
template <class F>
inline int guarded(F&& f) {
    try {
        f();
        return CFG_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host out of memory";
        return CFG_ERR_DEVICE;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return CFG_ERR_ARG;
    }
}

// largest dynamic LDS a block may opt into on this device (set once)
inline void init_device_limits() {
    static bool done = false;
    if (done) return;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    size_t b = p.sharedMemPerBlockOptin ? p.sharedMemPerBlockOptin : p.sharedMemPerBlock;
    set_lds_budget(std::min<size_t>(b, 160 * 1024));
    done = true;
}

// Device memory a hipMalloc of `bytes` takes (hipMemGetInfo deltas measured
// on the MI355X): more than 1 MiB rounds up to whole 2 MiB; smaller blocks are
// packed into 2 MiB chunks at up to 1.2x their size (1.25x counted here).
inline size_t device_bytes_taken(size_t bytes) {
    const size_t M1 = size_t(1) << 20, M2 = size_t(2) << 20;
    if (bytes > M1) return (bytes + M2 - 1) / M2 * M2;
    return (bytes + bytes / 4 + 4095) / 4096 * 4096;
}

// ---------------------------------------------------------------------------
// Device memory recycled by exact size. One pool serves one stream: a block
// released while kernels on that stream may still use it is only handed out
// again for work queued later on the same stream, so stream order keeps it
// safe. Blocks go back to HIP when the pool is destroyed or trimmed.
// ---------------------------------------------------------------------------
class DevPool {
   public:
    DevPool() = default;
    DevPool(const DevPool&) = delete;
    DevPool& operator=(const DevPool&) = delete;
    ~DevPool() {
        for (double* p : allocs_) (void)hipFree(p);
    }
    // A buffer of n doubles, reused from the free list of its size when it can
    // be. (Slabs of several buffers per hipMalloc made member set-up cheaper
    // but slowed the x-FFTs of a 16-member 64^3 batch by 15 %: one allocation
    // per buffer is kept.)
    double* take(size_t n) {
        auto& fl = free_[n];
        if (fl.empty()) {
            double* p = nullptr;
            HIPCHK(hipMalloc(&p, n * sizeof(double)));
            allocs_.insert(p);
            held_ += device_bytes_taken(n * sizeof(double));
            return p;
        }
        double* p = fl.back();
        fl.pop_back();
        return p;
    }
    void give(double* p, size_t n) {
        if (p) free_[n].push_back(p);
    }
    // device bytes of every block the pool holds (in use or free)
    size_t held() const { return held_; }
    // free blocks of n doubles, for a caller that plans allocations
    size_t free_count(size_t n) const {
        auto it = free_.find(n);
        return it == free_.end() ? 0 : it->second.size();
    }
    // returns the free blocks to HIP; the caller has synchronized the stream
    void trim() {
        for (auto& kv : free_) {
            for (double* p : kv.second) {
                (void)hipFree(p);
                allocs_.erase(p);
                held_ -= device_bytes_taken(kv.first * sizeof(double));
            }
        }
        free_.clear();
    }

   private:
    std::map<size_t, std::vector<double*>> free_;
    std::unordered_set<double*> allocs_;  // every allocation, freed with the pool
    size_t held_ = 0;
};

// Device copy of a small constant table, shared by every user with the same
// contents (the base-flow constants of the members of one grid, base flow and
// nu). Expired entries are swept as the cache grows.
inline std::shared_ptr<DBuf> shared_consts(const std::vector<double>& h) {
    static std::map<std::vector<double>, std::weak_ptr<DBuf>> cache;
    auto it = cache.find(h);
    if (it != cache.end()) {
        if (auto sp = it->second.lock()) return sp;
        cache.erase(it);
    }
    if (cache.size() >= 256)
        for (auto jt = cache.begin(); jt != cache.end();) jt = jt->second.expired() ? cache.erase(jt) : std::next(jt);
    auto sp = std::make_shared<DBuf>(h.size());
    HIPCHK(hipMemcpy(sp->p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    cache[h] = sp;
    return sp;
}

// A device buffer taken from a pool
struct PBuf {
    double* p = nullptr;
    size_t n = 0;
    DevPool* pool = nullptr;
    PBuf() = default;
    PBuf(DevPool& pl, size_t count) : n(count), pool(&pl) { p = count ? pl.take(count) : nullptr; }
    PBuf(const PBuf&) = delete;
    PBuf& operator=(const PBuf&) = delete;
    PBuf(PBuf&& o) noexcept : p(o.p), n(o.n), pool(o.pool) { o.p = nullptr; o.n = 0; }
    PBuf& operator=(PBuf&& o) noexcept {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(pool, o.pool);
        return *this;
    }
    ~PBuf() {
        if (p && pool) pool->give(p, n);
    }
    void zero(hipStream_t st) {
        if (n) HIPCHK(hipMemsetAsync(p, 0, n * sizeof(double), st));
    }
};

// ---------------------------------------------------------------------------
// Launch tables in device memory, keyed by their bytes. A table seen before
// is reused without a copy, so a periodic step sequence (the multistep
// history rotation) uploads nothing once warm. One cache serves one stream;
// when its arena fills, it waits for the stream and starts over.
// ---------------------------------------------------------------------------
// Arena bytes per cache: 4 MiB, or CFGPU_TABLE_ARENA_BYTES (at least 4 KiB) when
// set, which tests use to force recycles.
inline size_t table_arena_bytes() {
    const char* e = std::getenv("CFGPU_TABLE_ARENA_BYTES");
    if (e && *e) {
        const long v = std::atol(e);
        if (v >= 4096) return (size_t)v;
    }
    return size_t(4) << 20;
}

class TableCache {
   public:
    explicit TableCache(size_t bytes = table_arena_bytes()) : cap_(bytes) { arena_.alloc(bytes / sizeof(double)); }
    const void* get(const void* data, size_t n, hipStream_t st) {
        std::string key(static_cast<const char*>(data), n);
        auto it = map_.find(key);
        if (it != map_.end()) return it->second;
        const size_t need = (n + 255) & ~size_t(255);
        if (need > cap_) throw Error(CFG_ERR_ARG, "launch table larger than the table arena");
        if (used_ + need > cap_) {
            HIPCHK(hipStreamSynchronize(st));
            map_.clear();
            used_ = 0;
            ++gen_;
        }
        char* dst = reinterpret_cast<char*>(arena_.p) + used_;
        used_ += need;
        auto ins = map_.emplace(std::move(key), dst);
        // the key's bytes stay alive until the arena is recycled after a stream sync
        HIPCHK(hipMemcpyAsync(dst, ins.first->first.data(), n, hipMemcpyHostToDevice, st));
        return dst;
    }
    template <class T>
    const T* get(const std::vector<T>& v, hipStream_t st) {
        return static_cast<const T*>(get(v.data(), v.size() * sizeof(T), st));
    }
    // changes whenever the arena is recycled, i.e. when a table address may
    // start to hold other bytes
    unsigned long generation() const { return gen_; }

   private:
    DBuf arena_;
    size_t cap_, used_ = 0;
    unsigned long gen_ = 0;
    std::unordered_map<std::string, void*> map_;
};

// ---------------------------------------------------------------------------
// Per-grid constants shared by every field on that grid (process-wide)
// ---------------------------------------------------------------------------
struct GridKey {
    int Nx, Ny, Nz, dealias;
    double Lx, Lz, a, b;
    bool operator<(const GridKey& o) const {
        return std::tie(Nx, Ny, Nz, dealias, Lx, Lz, a, b) < std::tie(o.Nx, o.Ny, o.Nz, o.dealias, o.Lx, o.Lz, o.a, o.b);
    }
};
inline GridKey grid_key(const cfg_grid& g, bool dealias) {
    return GridKey{g.Nx, g.Ny, g.Nz, dealias ? 1 : 0, g.Lx, g.Lz, g.a, g.b};
}

inline std::shared_ptr<Geometry> geometry_for(const cfg_grid& gr, bool dealias) {
    static std::map<GridKey, std::weak_ptr<Geometry>> cache;
    const GridKey k = grid_key(gr, dealias);
    auto it = cache.find(k);
    if (it != cache.end())
        if (auto sp = it->second.lock()) return sp;
    auto geo = std::make_shared<Geometry>();
    geo->build(gr, dealias);
    cache[k] = geo;
    return geo;
}

// Device memory whose release waits: hipFree synchronises the whole device, so
// buffers that are dropped while other integrations run (tau data of a retired
// request) are collected here and freed together once enough has piled up.
inline size_t& deferred_pending_bytes() {
    static size_t total = 0;
    return total;
}
inline void deferred_free(double* p, size_t bytes, bool flush_all = false) {
    static std::vector<double*> pending;
    size_t& total = deferred_pending_bytes();
    if (p) {
        pending.push_back(p);
        total += bytes;
    }
    if (flush_all || total > (size_t(1) << 30)) {
        for (double* q : pending) (void)hipFree(q);
        pending.clear();
        total = 0;
    }
}


// Tau data of one (grid, nu, lambda_t, taucorrection), shared by every field
// that uses it (NSE::reset_lambda, nse.cpp:673-705, builds one per field).
// ready is recorded after the set-up kernel; users on any stream wait for it.
struct TauStore {
    double* coef = nullptr;
    double* scal = nullptr;
    size_t bytes = 0;
    TauSet set{};
    hipEvent_t ready = nullptr;
    TauStore() = default;
    TauStore(const TauStore&) = delete;
    TauStore& operator=(const TauStore&) = delete;
    ~TauStore() {
        if (ready) (void)hipEventDestroy(ready);
        deferred_free(coef, bytes);
        deferred_free(scal, 0);
    }
};

// device bytes of one tau set (one lambda) of a grid
inline size_t tau_set_bytes(const Geom& g) {
    const size_t prof = (size_t)g.Ny * g.M;
    return device_bytes_taken((size_t)TAU_NA * prof * sizeof(double)) +
           device_bytes_taken(6 * (size_t)g.M * sizeof(double));
}

inline std::shared_ptr<TauStore> tau_for(const Geometry& geo, double nu, double lambda_t, bool taucorr,
                                         hipStream_t st, DevPool& pool, KTimer* kt) {
    using Key = std::tuple<GridKey, double, double, int>;
    static std::map<Key, std::weak_ptr<TauStore>> cache;
    const Geom& g = geo.g;
    const GridKey gk{g.Nx, g.Ny, g.Nz, geo.dealias, g.Lx, g.Lz, g.a, g.b};
    const Key k{gk, nu, lambda_t, taucorr ? 1 : 0};
    auto it = cache.find(k);
    if (it != cache.end()) {
        if (auto sp = it->second.lock()) {
            HIPCHK(hipStreamWaitEvent(st, sp->ready, 0));
            return sp;
        }
        cache.erase(it);
    }
    auto ts = std::make_shared<TauStore>();
    const size_t prof = (size_t)g.Ny * g.M;
    HIPCHK(hipMalloc(&ts->coef, (size_t)TAU_NA * prof * sizeof(double)));
    HIPCHK(hipMalloc(&ts->scal, 6 * (size_t)g.M * sizeof(double)));
    ts->bytes = tau_set_bytes(g);
    ts->set.coef = ts->coef;
    ts->set.scal = ts->scal;
    ts->set.lambda_t = lambda_t;
    {
        hipEvent_t ev = kt ? kt->begin() : nullptr;
        launch_tau_setup(g, ts->set, nu, taucorr, st);
        if (kt) kt->end(K_SETUP, ev);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventCreateWithFlags(&ts->ready, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ts->ready, st));
    cache[k] = ts;
    return ts;
}

// Base flow U(y), W(y) of the flags as Chebyshev coefficients
// (NSE::createCFBaseFlow, nse.cpp:589-632).
inline void base_from_flags(const Geom& g, const cfg_flags& f, std::vector<double>& U, std::vector<double>& W) {
    const int Ny = g.Ny;
    U.assign(Ny, 0.0);
    W.assign(Ny, 0.0);
    switch (f.baseflow) {
        case CFG_ZeroBase: break;
        case CFG_LinearBase: U[1] = 1; break;
        case CFG_ParabolicBase: U[0] = 0.5; U[2] = -0.5; break;
        case CFG_SuctionBase:
            U = laminar_profile(f.nu, CFG_PressureGradient, 0, f.Ubulk, f.Vsuck, g.a, g.b, -0.5, 0.5, Ny);
            break;
        case CFG_LaminarBase:
            U = laminar_profile(f.nu, f.constraint, f.dPdx, f.Ubulk, f.Vsuck, g.a, g.b, f.ulowerwall, f.uupperwall, Ny);
            W = laminar_profile(f.nu, f.constraint, f.dPdz, f.Wbulk, f.Vsuck, g.a, g.b, f.wlowerwall, f.wupperwall, Ny);
            break;
        default: throw Error(CFG_ERR_ARG, "ArbitraryBase needs Ubase/Wbase");
    }
}

// The base flow as the nonlinear kernels read it: [4][Ny] = U, dU/dy, W, dW/dy
// at the y_n (NLMember::base).
inline std::vector<double> base_values_host(const Geom& g, const std::vector<long double>& tcheb,
                                            const std::vector<double>& U, const std::vector<double>& W) {
    std::vector<double> out;
    out.reserve(4 * g.Ny);
    for (const std::vector<double>* v : {&U, &W})
        for (bool d : {false, true}) {
            const std::vector<double> e = eval_profile(g, tcheb, *v, d);
            out.insert(out.end(), e.begin(), e.end());
        }
    return out;
}

// ---------------------------------------------------------------------------
// Flow constants of one field (nse.cpp:212-289, 589-671): base flow U(y),
// W(y), the constraint bookkeeping, and their device copies.
// ---------------------------------------------------------------------------
class FlowModel {
   public:
    FlowModel(const Geometry& geo, const cfg_flags& flags, const std::vector<double>& u0_packed, const double* Ubase,
              const double* Wbase)
        : flags_(flags) {
        const Geom& g = geo.g;
        const int Ny = g.Ny;
        if (Ubase) {
            Ubase_.assign(Ubase, Ubase + Ny);
            Wbase_.assign(Ny, 0.0);
            if (Wbase) Wbase_.assign(Wbase, Wbase + Ny);
        } else {
            create_base(g);
        }
        init_constraint(g, u0_packed);
        const double L = g.b - g.a;
        std::vector<double> Uyy = host_diff(host_diff(Ubase_, L), L), Wyy = host_diff(host_diff(Wbase_, L), L);
        std::vector<double> dev(6 * Ny);
        for (int i = 0; i < Ny; ++i) {
            dev[i] = flags_.nu * Uyy[i];
            dev[Ny + i] = flags_.nu * Wyy[i];
        }
        int off = 2 * Ny;
        for (auto* v : {&Ubase_, &Wbase_}) {  // values and y-derivatives at the Chebyshev points
            std::vector<double> val = geo.eval_profile(*v, false), der = geo.eval_profile(*v, true);
            std::copy(val.begin(), val.end(), dev.begin() + off);
            std::copy(der.begin(), der.end(), dev.begin() + off + Ny);
            off += 2 * Ny;
        }
        consts_ = shared_consts(dev);
        const std::vector<double> Uy = host_diff(Ubase_, L), Wy = host_diff(Wbase_, L);
        nuUslope_ = flags_.nu * (host_eval_b(Uy) - host_eval_a(Uy)) / L;
        nuWslope_ = flags_.nu * (host_eval_b(Wy) - host_eval_a(Wy)) / L;
        Ny_ = Ny;
    }

    const cfg_flags& flags() const { return flags_; }
    const std::vector<double>& Ubase() const { return Ubase_; }
    const std::vector<double>& Wbase() const { return Wbase_; }
    const double* base_values() const { return consts_->p + 2 * Ny_; }  // [4][Ny]: U, U', W, W' at y_n
    double UbulkAct() const { return UbulkAct_; }
    double dPdxAct0() const { return dPdxAct_; }
    double dPdzAct0() const { return dPdzAct_; }

    SolveConst solve_const(double* dP) const {
        return SolveConst{consts_->p, consts_->p + Ny_, dPdxRef_, dPdzRef_, UbulkRef_ - UbulkBase_, WbulkRef_ - WbulkBase_,
                          flags_.constraint == CFG_BulkVelocity ? 1 : 0, 0, dP, dP + 1};
    }
    LinearConst linear_const() const {
        return LinearConst{consts_->p, consts_->p + Ny_, dPdxRef_, dPdzRef_, flags_.constraint == CFG_BulkVelocity ? 1 : 0,
                           0, nuUslope_, nuWslope_};
    }

    void reset_gradp(double dPdx, double dPdz) {  // nse.cpp:723-734
        flags_.constraint = CFG_PressureGradient;
        flags_.dPdx = dPdx; flags_.dPdz = dPdz; flags_.Ubulk = 0.0; flags_.Wbulk = 0.0;
        dPdxRef_ = dPdx; dPdzRef_ = dPdz; UbulkRef_ = 0.0; WbulkRef_ = 0.0;
    }
    void reset_bulkv(double Ubulk, double Wbulk) {  // nse.cpp:735-746
        flags_.constraint = CFG_BulkVelocity;
        flags_.Ubulk = Ubulk; flags_.Wbulk = Wbulk; flags_.dPdx = 0.0; flags_.dPdz = 0.0;
        UbulkRef_ = Ubulk; WbulkRef_ = Wbulk; dPdxRef_ = 0.0; dPdzRef_ = 0.0;
    }

   private:
    void create_base(const Geom& g) { base_from_flags(g, flags_, Ubase_, Wbase_); }

    void init_constraint(const Geom& g, const std::vector<double>& up) {  // nse.cpp:634-671
        const int Ny = g.Ny;
        const double L = g.b - g.a;
        UbulkBase_ = host_mean(Ubase_);
        WbulkBase_ = host_mean(Wbase_);
        std::vector<double> u00(Ny), w00(Ny);
        for (int ny = 0; ny < Ny; ++ny) {  // real parts of the kx = kz = 0 profile (mode index 0)
            u00[ny] = up[2 * ((size_t)(0 * Ny + ny) * g.M)];
            w00[ny] = up[2 * ((size_t)(2 * Ny + ny) * g.M)];
        }
        auto bulk = [](double v) { return std::abs(v) < 1e-15 ? 0.0 : v; };
        UbulkAct_ = UbulkBase_ + bulk(host_mean(u00));
        WbulkAct_ = WbulkBase_ + bulk(host_mean(w00));
        std::vector<double> ut(Ny);
        for (int ny = 0; ny < Ny; ++ny) ut[ny] = u00[ny] + Ubase_[ny];
        std::vector<double> du = host_diff(ut, L), dw = host_diff(w00, L);
        dPdxAct_ = flags_.nu * (host_eval_b(du) - host_eval_a(du)) / L;
        dPdzAct_ = flags_.nu * (host_eval_b(dw) - host_eval_a(dw)) / L;
        if (flags_.constraint == CFG_BulkVelocity) {
            UbulkRef_ = flags_.Ubulk;
            WbulkRef_ = flags_.Wbulk;
        } else {
            dPdxAct_ = flags_.dPdx;
            dPdxRef_ = flags_.dPdx;
            dPdzAct_ = flags_.dPdz;
            dPdzRef_ = flags_.dPdz;
        }
    }

    cfg_flags flags_;
    int Ny_ = 0;
    std::vector<double> Ubase_, Wbase_;
    std::shared_ptr<DBuf> consts_;  // [6][Ny]: nu U'', nu W'' (coefficients), then U, U', W, W' at the y_n
    double nuUslope_ = 0, nuWslope_ = 0;
    double dPdxRef_ = 0, dPdxAct_ = 0, dPdzRef_ = 0, dPdzAct_ = 0, UbulkRef_ = 0, UbulkAct_ = 0, UbulkBase_ = 0;
    double WbulkRef_ = 0, WbulkAct_ = 0, WbulkBase_ = 0;
};

// Pinned host buffer that host <-> device field copies pass through (copies
// from pageable memory are staged by the runtime in small pieces). Grows as
// needed; one per process, used by one copy at a time.
inline double* pinned_stage(size_t n) {
    static double* buf = nullptr;
    static size_t cap = 0;
    if (n > cap) {
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        cap = 0;
        HIPCHK(hipHostMalloc((void**)&buf, n * sizeof(double), hipHostMallocDefault));
        cap = n;
    }
    return buf;
}

inline std::vector<double> upload_packed(const Geom& g, int nd, const double* ref, double* dst, hipStream_t st) {
    std::vector<double> h;
    pack(g, nd, ref, h);
    for (double v : h)
        if (!std::isfinite(v)) throw Error(CFG_ERR_NONFINITE, "non-finite value in the input field");
    double* pin = pinned_stage(h.size());
    std::memcpy(pin, h.data(), h.size() * sizeof(double));
    HIPCHK(hipMemcpyAsync(dst, pin, h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
}

inline void download_packed(const Geom& g, int nd, const double* src, size_t n, double* ref, hipStream_t st) {
    double* pin = pinned_stage(n);
    HIPCHK(hipMemcpyAsync(pin, src, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> h(pin, pin + n);
    unpack(g, nd, h, ref);
}

}  // namespace cfg
