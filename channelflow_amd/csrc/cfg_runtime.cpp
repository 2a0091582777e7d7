// Host runtime of the MI355X DNS time step: the DNS / NSE / time-stepping
// logic of channelflow (dns.cpp, dnsalgo.cpp, nse.cpp) driving the HIP
// kernels of cfg_kernels.hip, and the C ABI declared in include/cfgpu.h.
//
// One cfg_dns owns one HIP stream and all of its device buffers; the
// velocity, pressure and multistep history stay resident between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cfgpu.h"
#include "cfg_kernels.hpp"
#include "cfg_math.hpp"

namespace cfg {

static thread_local std::string g_err;

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
static long double cos_pi_frac(long num, long den) {  // cos(pi num/den), reduced exactly
    long r = num % (2 * den);
    if (r < 0) r += 2 * den;
    return cosl(3.141592653589793238462643383279502884L * (long double)r / (long double)den);
}

static std::vector<double> cheb_points(int N, double a, double b) {
    std::vector<double> x(N);
    const double piN = 3.14159265358979323846 / (N - 1);
    const double radius = (b - a) / 2, center = (b + a) / 2;
    for (int j = 0; j < N; ++j) x[j] = center + radius * cos(j * piN);
    return x;
}

// values at the Chebyshev points -> coefficients (the normalised DCT-I)
static std::vector<double> cheb_coeffs(const std::vector<double>& v) {
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

static std::vector<double> host_diff(const std::vector<double>& u, double L) {
    std::vector<double> du(u.size());
    std::vector<double> uu(u);
    cheb_diff(SV{uu.data(), 1}, SV{du.data(), 1}, (int)u.size(), L);
    return du;
}

static double host_eval_a(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_eval_a(SV{uu.data(), 1}, (int)u.size());
}
static double host_eval_b(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_eval_b(SV{uu.data(), 1}, (int)u.size());
}
static double host_mean(const std::vector<double>& u) {
    std::vector<double> uu(u);
    return cheb_mean(SV{uu.data(), 1}, (int)u.size());
}

// laminar profile (nse.cpp:823-943)
static std::vector<double> laminar_profile(double nu, int constraint, double dPdx, double Ubulk, double Vsuck,
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
struct FFTTable {
    FFTPlan plan{};
    DBuf tw;
    void build(int N) {
        int n = N, np = 0;
        while (n % 4 == 0) { plan.radix[np++] = 4; n /= 4; }
        while (n % 2 == 0) { plan.radix[np++] = 2; n /= 2; }
        while (n % 3 == 0) { plan.radix[np++] = 3; n /= 3; }
        if (n != 1) throw Error(CFG_ERR_ARG, "grid size " + std::to_string(N) + " is not of the form 2^a 3^b");
        plan.N = N;
        plan.npass = np;
        std::vector<double> h(2 * N);
        for (int j = 0; j < N; ++j) {
            const long double ang = -2.0L * 3.141592653589793238462643383279502884L * j / N;
            h[2 * j] = (double)cosl(ang);
            h[2 * j + 1] = (double)sinl(ang);
        }
        tw.alloc(2 * N);
        HIPCHK(hipMemcpy(tw.p, h.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice));
        plan.tw = reinterpret_cast<const double2*>(tw.p);
    }
};

struct Geometry {
    Geom g{};
    FFTTable fx, fz;
    DBuf T, TD, Ti, DTi;

    void build(const cfg_grid& gr, bool dealias) {
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
        fx.build(g.Nx);
        fz.build(g.Nz);
        const int N = g.Ny;
        const long double L = (long double)(gr.b - gr.a);
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
        std::vector<double> hT(N * N), hTD(N * N), hTi(N * N), hDTi(N * N);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                long double sd = 0.0L, sdi = 0.0L;
                for (int k = 0; k < N; ++k) {
                    sd += t[i * N + k] * d[k * N + j];
                    sdi += d[i * N + k] * ti[k * N + j];
                }
                hT[i * N + j] = (double)t[i * N + j];
                hTD[i * N + j] = (double)sd;
                hTi[i * N + j] = (double)ti[i * N + j];
                hDTi[i * N + j] = (double)sdi;
            }
        auto up = [&](DBuf& b, const std::vector<double>& h) {
            b.alloc(h.size());
            HIPCHK(hipMemcpy(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        };
        up(T, hT);
        up(TD, hTD);
        up(Ti, hTi);
        up(DTi, hDTi);
    }

    // y values of a coefficient profile and of its derivative (long double sums)
    std::vector<double> eval_profile(const std::vector<double>& c, bool deriv) const {
        const int N = g.Ny;
        std::vector<double> cd = deriv ? host_diff(c, g.b - g.a) : c;
        std::vector<double> v(N);
        for (int y = 0; y < N; ++y) {
            long double s = 0.0L;
            for (int k = 0; k < N; ++k) s += cos_pi_frac((long)y * k, N - 1) * (long double)cd[k];
            v[y] = (double)s;
        }
        return v;
    }
};

// compact <-> channelflow layout
static void pack(const Geom& g, int nd, const double* ref, std::vector<double>& out) {
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

static void unpack(const Geom& g, int nd, const std::vector<double>& in, double* ref) {
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

static int smooth235(int n) {
    for (int p : {2, 3})
        while (n % p == 0) n /= p;
    return n == 1;
}

static void check_grid(const cfg_grid& gr) {
    if (gr.Nx < 4 || gr.Nz < 4 || gr.Ny < 5 || gr.Ny % 2 == 0)
        throw Error(CFG_ERR_ARG, "grid must have Nx, Nz >= 4 and odd Ny >= 5 (Chebyshev-tau needs odd Ny)");
    if (gr.Nx % 2 || gr.Nz % 2) throw Error(CFG_ERR_ARG, "Nx and Nz must be even");
    if (!smooth235(gr.Nx) || !smooth235(gr.Nz)) throw Error(CFG_ERR_ARG, "Nx and Nz must be of the form 2^a 3^b");
    if (gr.Ny > 383) throw Error(CFG_ERR_ARG, "Ny > 383 is not supported by the tau-solve kernels");
    if (gr.Nx > 256 || gr.Nz > 512) throw Error(CFG_ERR_ARG, "Nx <= 256 and Nz <= 512 are supported");
    if (!(gr.b > gr.a) || !(gr.Lx > 0) || !(gr.Lz > 0)) throw Error(CFG_ERR_ARG, "invalid box");
}

static void check_flags(const cfg_flags& f) {
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
// shared device context of one DNS: stream, geometry, work buffers
// ---------------------------------------------------------------------------
struct Context {
    hipStream_t st = nullptr;
    Geometry geo;
    DBuf nlw_yp, nlw_xp, nlw_zo, nlw_fo, scratch;
    long nfield = 0;  // doubles in one 3-component compact field
    long nscal = 0;   // doubles in one 1-component compact field
    KTimer kt;

    Context(const cfg_grid& gr, bool dealias) {
        static bool lds_set = false;
        if (!lds_set) {  // largest dynamic LDS a block may opt into on this device
            int dev = 0;
            HIPCHK(hipGetDevice(&dev));
            hipDeviceProp_t p;
            HIPCHK(hipGetDeviceProperties(&p, dev));
            size_t b = p.sharedMemPerBlockOptin ? p.sharedMemPerBlockOptin : p.sharedMemPerBlock;
            set_lds_budget(std::min<size_t>(b, 160 * 1024));
            lds_set = true;
        }
        HIPCHK(hipStreamCreate(&st));
        kt.st = st;
        geo.build(gr, dealias);
        const Geom& g = geo.g;
        nfield = 2L * 3 * g.Ny * g.M;
        nscal = 2L * g.Ny * g.M;
        nlw_yp.alloc(2L * 6 * g.Ny * g.M);
        nlw_xp.alloc(2L * 9 * g.Ny * g.Nx * g.nkz);
        nlw_zo.alloc(2L * 9 * g.Ny * g.Nx * g.nkz);
        nlw_fo.alloc(2L * 9 * g.Ny * g.M);
        scratch.alloc(4L * g.Ny * g.M);
    }
    ~Context() {
        if (st) (void)hipStreamDestroy(st);
    }
    NLWork work() { return NLWork{nlw_yp.p, nlw_xp.p, nlw_zo.p, nlw_fo.p}; }
};

// A velocity/pressure pair, compact layout
struct Fields {
    DBuf u, q;
    void alloc(const Context& c) {
        u.alloc(c.nfield);
        q.alloc(c.nscal);
    }
    void zero(hipStream_t st) {
        u.zero(st);
        q.zero(st);
    }
    void copy_from(const Fields& o, hipStream_t st) {
        HIPCHK(hipMemcpyAsync(u.p, o.u.p, u.n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(q.p, o.q.p, q.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
};

static void swap_fields(Fields& a, Fields& b) {
    std::swap(a.u, b.u);
    std::swap(a.q, b.q);
}

// ---------------------------------------------------------------------------
// NSE (nse.cpp): base flow, constraint bookkeeping, nonlinear/linear/solve
// ---------------------------------------------------------------------------
struct TauStore {
    DBuf coef, scal;
    TauSet set{};
};

class NSE {
   public:
    NSE(Context& ctx, const cfg_flags& flags, const std::vector<double>& u0_packed, const double* Ubase,
        const double* Wbase)
        : c_(ctx), flags_(flags) {
        const Geom& g = c_.geo.g;
        const int Ny = g.Ny;
        if (Ubase) {
            Ubase_.assign(Ubase, Ubase + Ny);
            Wbase_.assign(Wbase ? Wbase : Ubase, (Wbase ? Wbase : Ubase) + Ny);
            if (!Wbase) std::fill(Wbase_.begin(), Wbase_.end(), 0.0);
        } else {
            create_base();
        }
        init_constraint(u0_packed);
        const double L = g.b - g.a;
        std::vector<double> Uyy = host_diff(host_diff(Ubase_, L), L), Wyy = host_diff(host_diff(Wbase_, L), L);
        std::vector<double> nuU(Ny), nuW(Ny);
        for (int i = 0; i < Ny; ++i) {
            nuU[i] = flags_.nu * Uyy[i];
            nuW[i] = flags_.nu * Wyy[i];
        }
        nuUyy_.alloc(Ny);
        nuWyy_.alloc(Ny);
        HIPCHK(hipMemcpy(nuUyy_.p, nuU.data(), Ny * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(nuWyy_.p, nuW.data(), Ny * sizeof(double), hipMemcpyHostToDevice));
        std::vector<double> b4;
        for (auto* v : {&Ubase_, &Wbase_}) {
            std::vector<double> val = c_.geo.eval_profile(*v, false), der = c_.geo.eval_profile(*v, true);
            b4.insert(b4.end(), val.begin(), val.end());
            b4.insert(b4.end(), der.begin(), der.end());
        }
        base_.alloc(4 * Ny);
        HIPCHK(hipMemcpy(base_.p, b4.data(), 4 * Ny * sizeof(double), hipMemcpyHostToDevice));
        const std::vector<double> Uy = host_diff(Ubase_, L), Wy = host_diff(Wbase_, L);
        nuUslope_ = flags_.nu * (host_eval_b(Uy) - host_eval_a(Uy)) / L;
        nuWslope_ = flags_.nu * (host_eval_b(Wy) - host_eval_a(Wy)) / L;
        dP_.alloc(2);
        std::vector<double> dp = {dPdxAct_, dPdzAct_};
        HIPCHK(hipMemcpy(dP_.p, dp.data(), 2 * sizeof(double), hipMemcpyHostToDevice));
        method_ = flags_.nonlinearity;
    }

    void reset_lambda(const std::vector<double>& lambda_t) {  // nse.cpp:673-705
        const Geom& g = c_.geo.g;
        const size_t prof = (size_t)g.Ny * g.M;
        tau_.clear();
        tau_.resize(lambda_t.size());
        for (size_t j = 0; j < lambda_t.size(); ++j) {
            TauStore& ts = tau_[j];
            ts.coef.alloc((size_t)TAU_NA * prof);
            ts.scal.alloc(6 * (size_t)g.M);
            ts.set.coef = ts.coef.p;
            ts.set.scal = ts.scal.p;
            ts.set.lambda_t = lambda_t[j];
            hipEvent_t ev = c_.kt.begin();
            launch_tau_setup(g, ts.set, flags_.nu, flags_.taucorrection != 0, c_.scratch.p, c_.st);
            c_.kt.end(K_SETUP, ev);
        }
        HIPCHK(hipGetLastError());
    }

    // NSE::nonlinear (nse.cpp:383-391)
    void nonlinear(const double* u, double* f) {
        int m;
        switch (method_) {
            case CFG_Rotational: m = NL_ROT; break;
            case CFG_Convection: m = NL_CONV; break;
            case CFG_Divergence: m = NL_DIV; break;
            case CFG_SkewSymmetric: m = NL_SKEW; break;
            case CFG_Alternating: m = NL_DIV; method_ = CFG_Alternating_; break;
            default: m = NL_CONV; method_ = CFG_Alternating; break;
        }
        YMats ym{c_.geo.T.p, c_.geo.TD.p, c_.geo.Ti.p, c_.geo.DTi.p, base_.p};
        launch_nonlinear(c_.geo.g, ym, c_.geo.fx.plan, c_.geo.fz.plan, m, flags_.Vsuck, flags_.rotation, u, f,
                         c_.work(), c_.st, &c_.kt);
        HIPCHK(hipGetLastError());
    }

    // NSE::linear (nse.cpp:393-477)
    void linear(const double* u, const double* q, double* out) {
        LinearConst lc{nuUyy_.p, nuWyy_.p, dPdxRef_, dPdzRef_, flags_.constraint == CFG_BulkVelocity ? 1 : 0,
                       nuUslope_, nuWslope_};
        hipEvent_t ev = c_.kt.begin();
        launch_linear(c_.geo.g, flags_.nu, lc, u, q, out, c_.scratch.p, c_.st);
        c_.kt.end(K_LINEAR, ev);
        HIPCHK(hipGetLastError());
    }

    // NSE::solve (nse.cpp:479-575). rhs is consumed (kx=kz=0 entries modified).
    void solve(double* u, double* q, double* rhs, int s) {
        SolveConst sc{nuUyy_.p, nuWyy_.p, dPdxRef_, dPdzRef_, UbulkRef_ - UbulkBase_, WbulkRef_ - WbulkBase_,
                      flags_.constraint == CFG_BulkVelocity ? 1 : 0, dP_.p, dP_.p + 1};
        hipEvent_t ev = c_.kt.begin();
        launch_tau_solve(c_.geo.g, tau_.at(s).set, flags_.nu, flags_.taucorrection != 0, sc, rhs, u, q,
                         c_.scratch.p, c_.st);
        c_.kt.end(K_TAU, ev);
        HIPCHK(hipGetLastError());
        if (flags_.constraint == CFG_BulkVelocity) dP_dirty_ = true;
    }

    double dPdx() {
        if (flags_.constraint == CFG_BulkVelocity && dP_dirty_) {
            double h[2];
            HIPCHK(hipMemcpyAsync(h, dP_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c_.st));
            HIPCHK(hipStreamSynchronize(c_.st));
            dPdxAct_ = h[0];
            dPdzAct_ = h[1];
            dP_dirty_ = false;
        }
        return dPdxAct_;
    }
    double Ubulk() const { return UbulkAct_; }
    const std::vector<double>& Ubase() const { return Ubase_; }
    const std::vector<double>& Wbase() const { return Wbase_; }
    const cfg_flags& flags() const { return flags_; }

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
    void create_base() {  // nse.cpp:589-632
        const Geom& g = c_.geo.g;
        const int Ny = g.Ny;
        Ubase_.assign(Ny, 0.0);
        Wbase_.assign(Ny, 0.0);
        switch (flags_.baseflow) {
            case CFG_ZeroBase: break;
            case CFG_LinearBase: Ubase_[1] = 1; break;
            case CFG_ParabolicBase: Ubase_[0] = 0.5; Ubase_[2] = -0.5; break;
            case CFG_SuctionBase:
                Ubase_ = laminar_profile(flags_.nu, CFG_PressureGradient, 0, flags_.Ubulk, flags_.Vsuck, g.a, g.b, -0.5,
                                         0.5, Ny);
                break;
            case CFG_LaminarBase:
                Ubase_ = laminar_profile(flags_.nu, flags_.constraint, flags_.dPdx, flags_.Ubulk, flags_.Vsuck, g.a, g.b,
                                         flags_.ulowerwall, flags_.uupperwall, Ny);
                Wbase_ = laminar_profile(flags_.nu, flags_.constraint, flags_.dPdz, flags_.Wbulk, flags_.Vsuck, g.a, g.b,
                                         flags_.wlowerwall, flags_.wupperwall, Ny);
                break;
            default: throw Error(CFG_ERR_ARG, "ArbitraryBase needs Ubase/Wbase");
        }
    }

    void init_constraint(const std::vector<double>& up) {  // nse.cpp:634-671
        const Geom& g = c_.geo.g;
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

    Context& c_;
    cfg_flags flags_;
    int method_ = 0;
    std::vector<double> Ubase_, Wbase_;
    DBuf nuUyy_, nuWyy_, base_, dP_;
    double nuUslope_ = 0, nuWslope_ = 0;
    double dPdxRef_ = 0, dPdxAct_ = 0, dPdzRef_ = 0, dPdzAct_ = 0, UbulkRef_ = 0, UbulkAct_ = 0, UbulkBase_ = 0;
    double WbulkRef_ = 0, WbulkAct_ = 0, WbulkBase_ = 0;
    bool dP_dirty_ = false;
    std::vector<TauStore> tau_;
};

// ---------------------------------------------------------------------------
// time-stepping algorithms (dnsalgo.cpp)
// ---------------------------------------------------------------------------
class Algorithm {
   public:
    Algorithm(Context& c, std::shared_ptr<NSE> nse, const cfg_flags& f) : c_(c), nse_(std::move(nse)), flags_(f), t_(f.t0) {}
    virtual ~Algorithm() = default;
    virtual void advance(Fields& cur, int n) = 0;
    virtual void push(Fields& cur) {}
    virtual bool full() const { return true; }
    virtual void reset_dt(double dt) = 0;
    virtual int Ninitsteps() const { return 0; }
    double time() const { return t_; }
    void reset_time(double t) { t_ = t; }
    double dt() const { return flags_.dt; }

   protected:
    void combine(double* out, std::initializer_list<std::pair<double, const double*>> terms) {
        double coef[8];
        const double* in[8];
        int n = 0;
        for (auto& t : terms) {
            coef[n] = t.first;
            in[n] = t.second;
            ++n;
        }
        hipEvent_t ev = c_.kt.begin();
        launch_combine(c_.nfield, n, coef, in, out, c_.st);
        c_.kt.end(K_COMBINE, ev);
    }
    Context& c_;
    std::shared_ptr<NSE> nse_;
    cfg_flags flags_;
    double t_;
};

class Multistep : public Algorithm {  // dnsalgo.cpp:102-315
   public:
    Multistep(Context& c, std::shared_ptr<NSE> nse, const cfg_flags& f) : Algorithm(c, std::move(nse), f) {
        switch (f.timestepping) {
            case CFG_CNFE1:
            case CFG_SBDF1: order_ = 1; eta_ = 1.0; alpha_ = {-1.0}; beta_ = {1.0}; break;
            case CFG_SBDF2: order_ = 2; eta_ = 1.5; alpha_ = {-2.0, 0.5}; beta_ = {2.0, -1.0}; break;
            case CFG_SBDF3:
                order_ = 3; eta_ = 11.0 / 6.0; alpha_ = {-3.0, 1.5, -1.0 / 3.0}; beta_ = {3.0, -3.0, 1.0};
                break;
            default:
                order_ = 4; eta_ = 25.0 / 12.0; alpha_ = {-4.0, 3.0, -4.0 / 3.0, 0.25}; beta_ = {4.0, -6.0, 4.0, -1.0};
        }
        nse_->reset_lambda({eta_ / flags_.dt});
        fields_.resize(order_);
        nonlf_.resize(order_);
        for (int j = 0; j < order_; ++j) {
            fields_[j].alloc(c_);
            fields_[j].zero(c_.st);
            nonlf_[j].alloc(c_.nfield);
            nonlf_[j].zero(c_.st);
        }
        rhs_.alloc(c_.nfield);
        countdown_ = order_ - 1;
    }
    bool full() const override { return countdown_ == 0; }
    int Ninitsteps() const override { return order_ - 1; }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        nse_->reset_lambda({eta_ / flags_.dt});
        countdown_ = order_ - 1;
    }
    void push(Fields& cur) override {
        rotate();
        if (order_ > 1) {
            fields_[1].copy_from(cur, c_.st);
            nse_->nonlinear(fields_[1].u.p, nonlf_[1].p);
        }
        t_ += flags_.dt;
        --countdown_;
    }
    void advance(Fields& cur, int N) override {
        const int J = order_ - 1;
        fields_[0].copy_from(cur, c_.st);
        for (int step = 0; step < N; ++step) {
            nse_->nonlinear(fields_[0].u.p, nonlf_[0].p);
            double coef[8];
            const double* in[8];
            int n = 0;
            for (int j = 0; j < order_; ++j) {
                coef[n] = -alpha_[j] / flags_.dt;
                in[n++] = fields_[j].u.p;
                coef[n] = -beta_[j];
                in[n++] = nonlf_[j].p;
            }
            hipEvent_t ev = c_.kt.begin();
            launch_combine(c_.nfield, n, coef, in, rhs_.p, c_.st);
            c_.kt.end(K_COMBINE, ev);
            nse_->solve(fields_[J].u.p, fields_[J].q.p, rhs_.p, 0);
            rotate();
            t_ += flags_.dt;
        }
        cur.copy_from(fields_[0], c_.st);
    }

   private:
    void rotate() {
        for (int j = order_ - 1; j > 0; --j) {
            std::swap(nonlf_[j], nonlf_[j - 1]);
            swap_fields(fields_[j], fields_[j - 1]);
        }
    }
    int order_ = 0, countdown_ = 0;
    double eta_ = 0;
    std::vector<double> alpha_, beta_;
    std::vector<Fields> fields_;
    std::vector<DBuf> nonlf_;
    DBuf rhs_;
};

class CNABstyle : public Algorithm {  // dnsalgo.cpp:510-697
   public:
    CNABstyle(Context& c, std::shared_ptr<NSE> nse, const cfg_flags& f) : Algorithm(c, std::move(nse), f) {
        if (f.timestepping == CFG_CNAB2) {
            nsub_ = 1; ninit_ = 1; full_ = false;
            alpha_ = {0.5}; beta_ = {0.5}; gamma_ = {1.5}; zeta_ = {-0.5};
        } else {
            nsub_ = 3; ninit_ = 0; full_ = true;
            alpha_ = {29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0};
            beta_ = {37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0};
            gamma_ = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
            zeta_ = {0.0, -17.0 / 60.0, -5.0 / 12.0};
        }
        for (auto* b : {&fj_, &fj1_, &lt_, &rhs_}) {
            b->alloc(c_.nfield);
            b->zero(c_.st);
        }
        set_lambda();
    }
    bool full() const override { return full_; }
    int Ninitsteps() const override { return ninit_; }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        set_lambda();
        if (flags_.timestepping == CFG_CNAB2) full_ = false;
    }
    void push(Fields& cur) override {
        std::swap(fj_, fj1_);
        nse_->nonlinear(cur.u.p, fj_.p);
        t_ += flags_.dt;
        full_ = true;
    }
    void advance(Fields& cur, int N) override {
        for (int n = 0; n < N; ++n) {
            for (int j = 0; j < nsub_; ++j) {
                std::swap(fj_, fj1_);
                nse_->nonlinear(cur.u.p, fj_.p);
                const double a_b = alpha_[j] / beta_[j], g_b = gamma_[j] / beta_[j], z_b = zeta_[j] / beta_[j];
                nse_->linear(cur.u.p, cur.q.p, lt_.p);
                combine(rhs_.p, {{lambda_[j], cur.u.p}, {a_b, lt_.p}, {-g_b, fj_.p}, {-z_b, fj1_.p}});
                nse_->solve(cur.u.p, cur.q.p, rhs_.p, j);
            }
            t_ += flags_.dt;
        }
    }

   private:
    void set_lambda() {
        lambda_.resize(nsub_);
        for (int j = 0; j < nsub_; ++j) lambda_[j] = 1.0 / (beta_[j] * flags_.dt);
        nse_->reset_lambda(lambda_);
    }
    int nsub_ = 0, ninit_ = 0;
    bool full_ = false;
    std::vector<double> alpha_, beta_, gamma_, zeta_, lambda_;
    DBuf fj_, fj1_, lt_, rhs_;
};

class RungeKutta : public Algorithm {  // dnsalgo.cpp:351-459
   public:
    RungeKutta(Context& c, std::shared_ptr<NSE> nse, const cfg_flags& f) : Algorithm(c, std::move(nse), f) {
        A_ = {0.0, -5.0 / 9.0, -153.0 / 128.0};
        B_ = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
        C_ = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
        for (auto* b : {&Qj_, &Qj1_, &lt_, &rhs_}) {
            b->alloc(c_.nfield);
            b->zero(c_.st);
        }
        set_lambda();
    }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        set_lambda();
    }
    void advance(Fields& cur, int N) override {
        for (int n = 0; n < N; ++n) {
            for (int j = 0; j < 3; ++j) {
                nse_->nonlinear(cur.u.p, Qj1_.p);
                combine(Qj_.p, {{A_[j], Qj_.p}, {-1.0, Qj1_.p}});
                nse_->linear(cur.u.p, cur.q.p, lt_.p);
                combine(rhs_.p, {{1.0, lt_.p}, {lambda_[j], cur.u.p}, {B_[j] / C_[j], Qj_.p}});
                nse_->solve(cur.u.p, cur.q.p, rhs_.p, j);
            }
            t_ += flags_.dt;
        }
    }

   private:
    void set_lambda() {
        lambda_.resize(3);
        for (int j = 0; j < 3; ++j) lambda_[j] = 1.0 / (C_[j] * flags_.dt);
        nse_->reset_lambda(lambda_);
    }
    std::vector<double> A_, B_, C_, lambda_;
    DBuf Qj_, Qj1_, lt_, rhs_;
};

static std::unique_ptr<Algorithm> new_algorithm(Context& c, std::shared_ptr<NSE> nse, const cfg_flags& f) {
    switch (f.timestepping) {
        case CFG_CNFE1:
        case CFG_SBDF1:
        case CFG_SBDF2:
        case CFG_SBDF3:
        case CFG_SBDF4: return std::make_unique<Multistep>(c, nse, f);
        case CFG_CNRK2: return std::make_unique<RungeKutta>(c, nse, f);
        default: return std::make_unique<CNABstyle>(c, nse, f);
    }
}

}  // namespace cfg

using namespace cfg;

// ---------------------------------------------------------------------------
// DNS (dns.cpp)
// ---------------------------------------------------------------------------
struct cfg_dns {
    std::unique_ptr<Context> ctx;
    cfg_flags flags{};
    cfg_grid grid{};
    Fields cur;
    std::shared_ptr<NSE> main_nse, init_nse;
    std::unique_ptr<Algorithm> main_alg, init_alg;

    void advance(int N) {  // dns.cpp:135-163
        if (!main_alg->full() && !init_alg) throw Error(CFG_ERR_STATE, "uninitialised multistep algorithm");
        int n = 0;
        while (!main_alg->full() && n < N) {
            main_alg->push(cur);
            init_alg->advance(cur, 1);
            ++n;
        }
        main_alg->advance(cur, N - n);
    }
};

template <class F>
static int guarded(F&& f) {
    try {
        f();
        return CFG_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CFG_ERR_ARG;
    }
}

static void upload(const Geom& g, int nd, const double* ref, DBuf& dst, hipStream_t st, std::vector<double>* keep) {
    std::vector<double> h;
    pack(g, nd, ref, h);
    for (double v : h)
        if (!std::isfinite(v)) throw Error(CFG_ERR_NONFINITE, "non-finite value in the input field");
    HIPCHK(hipMemcpyAsync(dst.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    if (keep) *keep = std::move(h);
}

static void download(const Geom& g, int nd, const DBuf& src, double* ref, hipStream_t st) {
    std::vector<double> h(src.n);
    HIPCHK(hipMemcpyAsync(h.data(), src.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    unpack(g, nd, h, ref);
}

extern "C" {

const char* cfg_last_error(void) { return g_err.c_str(); }

int cfg_device_count(int* n) {
    return guarded([&] { HIPCHK(hipGetDeviceCount(n)); });
}

int cfg_dns_create(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, const double* q_spec,
                   const double* Ubase, const double* Wbase, cfg_dns** out) {
    return guarded([&] {
        if (!grid || !flags || !u_spec || !out) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        auto d = std::make_unique<cfg_dns>();
        d->flags = *flags;
        d->grid = *grid;
        d->ctx = std::make_unique<Context>(*grid, flags->dealiasing == CFG_DealiasXZ);
        Context& c = *d->ctx;
        d->cur.alloc(c);
        std::vector<double> up;
        upload(c.geo.g, 3, u_spec, d->cur.u, c.st, &up);
        if (q_spec) upload(c.geo.g, 1, q_spec, d->cur.q, c.st, nullptr);
        else d->cur.q.zero(c.st);
        cfg_flags f = *flags;
        if (Ubase) f.baseflow = CFG_ArbitraryBase;
        d->main_nse = std::make_shared<NSE>(c, f, up, Ubase, Wbase);
        d->main_alg = new_algorithm(c, d->main_nse, f);
        if (!d->main_alg->full() && f.initstepping != f.timestepping) {  // dns.cpp:26-38
            cfg_flags fi = f;
            fi.timestepping = f.initstepping;
            d->init_nse = std::make_shared<NSE>(c, fi, up, Ubase, Wbase);
            d->init_alg = new_algorithm(c, d->init_nse, fi);
            if (d->init_alg->Ninitsteps() != 0)
                throw Error(CFG_ERR_ARG, "initstepping method needs initialisation itself");
        }
        HIPCHK(hipStreamSynchronize(c.st));
        *out = d.release();
    });
}

int cfg_dns_destroy(cfg_dns* dns) {
    return guarded([&] {
        if (dns) {
            (void)hipStreamSynchronize(dns->ctx->st);
            delete dns;
        }
    });
}

int cfg_dns_advance(cfg_dns* dns, int nsteps) {
    return guarded([&] {
        if (!dns || nsteps < 0) throw Error(CFG_ERR_ARG, "bad argument");
        dns->advance(nsteps);
        HIPCHK(hipStreamSynchronize(dns->ctx->st));
    });
}

int cfg_dns_set_state(cfg_dns* dns, const double* u_spec, const double* q_spec) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        const Geom& g = dns->ctx->geo.g;
        if (u_spec) upload(g, 3, u_spec, dns->cur.u, dns->ctx->st, nullptr);
        if (q_spec) upload(g, 1, q_spec, dns->cur.q, dns->ctx->st, nullptr);
    });
}

int cfg_dns_get_state(cfg_dns* dns, double* u_spec, double* q_spec) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        const Geom& g = dns->ctx->geo.g;
        if (u_spec) download(g, 3, dns->cur.u, u_spec, dns->ctx->st);
        if (q_spec) download(g, 1, dns->cur.q, q_spec, dns->ctx->st);
    });
}

int cfg_dns_reset_dt(cfg_dns* dns, double dt) {  // dns.cpp:180-185
    return guarded([&] {
        if (!dns || !(dt > 0)) throw Error(CFG_ERR_ARG, "bad argument");
        dns->main_alg->reset_dt(dt);
        if (dns->init_alg) dns->init_alg->reset_dt(dt);
        dns->flags.dt = dt;
        HIPCHK(hipStreamSynchronize(dns->ctx->st));
    });
}

int cfg_dns_reset_time(cfg_dns* dns, double t) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->main_alg->reset_time(t);
        if (dns->init_alg) dns->init_alg->reset_time(t);
    });
}

int cfg_dns_reset_gradp(cfg_dns* dns, double dPdx, double dPdz) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        if (dns->init_nse) dns->init_nse->reset_gradp(dPdx, dPdz);
        dns->main_nse->reset_gradp(dPdx, dPdz);
    });
}

int cfg_dns_reset_bulkv(cfg_dns* dns, double Ubulk, double Wbulk) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        if (dns->init_nse) dns->init_nse->reset_bulkv(Ubulk, Wbulk);
        dns->main_nse->reset_bulkv(Ubulk, Wbulk);
    });
}

int cfg_dns_time(const cfg_dns* dns, double* t) {
    return guarded([&] {
        if (!dns || !t) throw Error(CFG_ERR_ARG, "null argument");
        *t = dns->main_alg->time();
    });
}

int cfg_dns_dpdx(cfg_dns* dns, double* v) {
    return guarded([&] {
        if (!dns || !v) throw Error(CFG_ERR_ARG, "null argument");
        *v = dns->main_nse->dPdx();
    });
}

int cfg_dns_ubulk(const cfg_dns* dns, double* v) {
    return guarded([&] {
        if (!dns || !v) throw Error(CFG_ERR_ARG, "null argument");
        *v = dns->main_nse->Ubulk();
    });
}

int cfg_dns_cfl(cfg_dns* dns, double* cfl) {
    return guarded([&] {
        if (!dns || !cfl) throw Error(CFG_ERR_ARG, "null argument");
        throw Error(CFG_ERR_STATE, "CFL is not implemented on the device path yet");
    });
}

int cfg_dns_base(const cfg_dns* dns, double* U, double* W) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        const auto& Ub = dns->main_nse->Ubase();
        const auto& Wb = dns->main_nse->Wbase();
        if (U) std::copy(Ub.begin(), Ub.end(), U);
        if (W) std::copy(Wb.begin(), Wb.end(), W);
    });
}

int cfg_dns_device_state(cfg_dns* dns, double** u_dev, double** q_dev, int* nkx, int* nkz) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        if (u_dev) *u_dev = dns->cur.u.p;
        if (q_dev) *q_dev = dns->cur.q.p;
        if (nkx) *nkx = dns->ctx->geo.g.nkx;
        if (nkz) *nkz = dns->ctx->geo.g.nkz;
    });
}

int cfg_dns_profile(cfg_dns* dns, int enable) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->ctx->kt.reset();
        dns->ctx->kt.on = enable != 0;
    });
}

int cfg_dns_kernel_times(cfg_dns* dns, double* total_ms, long long* counts) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->ctx->kt.flush();
        for (int i = 0; i < K_COUNT; ++i) {
            if (total_ms) total_ms[i] = dns->ctx->kt.total_ms[i];
            if (counts) counts[i] = dns->ctx->kt.count[i];
        }
    });
}

const char* cfg_kernel_name(int id) {
    static const char* names[K_COUNT] = {"k_yinv", "k_xinv", "k_zpass", "k_xfwd", "k_yfwd",
                                         "k_combine", "k_tau_solve", "k_linear", "k_tau_setup"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "";
}

int cfg_nonlinear(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, double* f_spec) {
    return guarded([&] {
        if (!grid || !flags || !u_spec || !f_spec) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        Context c(*grid, flags->dealiasing == CFG_DealiasXZ);
        Fields u;
        u.alloc(c);
        std::vector<double> up;
        upload(c.geo.g, 3, u_spec, u.u, c.st, &up);
        NSE nse(c, *flags, up, nullptr, nullptr);
        DBuf f(c.nfield);
        nse.nonlinear(u.u.p, f.p);
        download(c.geo.g, 3, f, f_spec, c.st);
    });
}

int cfg_solve(const cfg_grid* grid, const cfg_flags* flags, double lambda_t, const double* u0_spec,
              const double* rhs_spec, double* u_spec, double* q_spec, double* dPdx) {
    return guarded([&] {
        if (!grid || !flags || !u0_spec || !rhs_spec || !u_spec || !q_spec) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        Context c(*grid, flags->dealiasing == CFG_DealiasXZ);
        const Geom& g = c.geo.g;
        Fields u;
        u.alloc(c);
        std::vector<double> up;
        upload(g, 3, u0_spec, u.u, c.st, &up);
        NSE nse(c, *flags, up, nullptr, nullptr);
        nse.reset_lambda({lambda_t});
        DBuf rhs(c.nfield);
        upload(g, 3, rhs_spec, rhs, c.st, nullptr);
        nse.solve(u.u.p, u.q.p, rhs.p, 0);
        // live modes overwritten, the rest of the caller's buffers untouched
        std::vector<double> hu(c.nfield), hq(c.nscal);
        HIPCHK(hipMemcpyAsync(hu.data(), u.u.p, hu.size() * sizeof(double), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipMemcpyAsync(hq.data(), u.q.p, hq.size() * sizeof(double), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        const int Mz = g.Nz / 2 + 1;
        for (int comp = 0; comp < 4; ++comp) {
            const std::vector<double>& src = comp < 3 ? hu : hq;
            double* dst = comp < 3 ? u_spec : q_spec;
            const int cc = comp < 3 ? comp : 0;
            for (int ny = 0; ny < g.Ny; ++ny)
                for (int ikz = 0; ikz < g.nkz; ++ikz)
                    for (int ikx = 0; ikx < g.nkx; ++ikx) {
                        const int kx = ikx <= g.kxd ? ikx : ikx - g.nkx;
                        const int mx = kx >= 0 ? kx : kx + g.Nx;
                        const size_t d = (size_t)ikz + Mz * ((size_t)mx + g.Nx * ((size_t)ny + g.Ny * cc));
                        const size_t s = (((size_t)cc * g.Ny + ny) * g.nkz + ikz) * g.nkx + ikx;
                        dst[2 * d] = src[2 * s];
                        dst[2 * d + 1] = src[2 * s + 1];
                    }
        }
        if (dPdx) *dPdx = nse.dPdx();
    });
}

}  // extern "C"
