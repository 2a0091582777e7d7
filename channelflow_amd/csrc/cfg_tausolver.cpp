// C ABI of channelflow's per-mode TauSolver (channelflow/tausolver.h:33-66):
// the constructor, solve, the bulk-velocity solve and verify, on the device.
//
// A cfg_tausolver owns one mode's tau data (the two factored Helmholtz
// operators, the homogeneous and tau-correction solutions, the influence
// matrix: the layout of one mode of a TauSet) and a small device work area.
// Each call runs one single-lane kernel over the per-mode routines of
// cfg_math.hpp, the same routines the batched set-up of the DNS uses. The
// DNS's own solves batch all modes in wave-parallel kernels; this object is
// the per-mode interface for callers that build TauSolvers themselves.
#include "cfg_host.hpp"

using namespace cfg;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

// io layout (doubles): Rx, Ry, Rz, u, v, w, P as interleaved complex
// Ny-vectors, then the two results of the bulk solve, then 4 real scratch
// Ny-vectors, then the verify result
struct TSLayout {
    long Ny;
    __host__ __device__ long cvec(int k) const { return 2 * Ny * k; }
    __host__ __device__ long dP() const { return 14 * Ny; }
    __host__ __device__ long scratch(int k) const { return 14 * Ny + 2 + Ny * k; }
    __host__ __device__ long err() const { return 18 * Ny + 2; }
    __host__ __device__ long total() const { return 18 * Ny + 3; }
};

__device__ inline CV cv_of(double* p) { return CV{SV{p, 2}, SV{p + 1, 2}}; }

struct TSParams {
    TauView t;
    double* io;
    long Ny;
};

__global__ void k_ts_setup(TSParams a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const TSLayout L{a.Ny};
    tau_setup(a.t, SV{a.io + L.scratch(0), 1}, SV{a.io + L.scratch(1), 1});
}

// TauSolver::solve (tausolver.cpp:347-402), or the bulk form (tausolver.cpp:404-450)
__global__ void k_ts_solve(TSParams a, int bulk, double umean, double wmean) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const TSLayout L{a.Ny};
    double* io = a.io;
    const CV Rx = cv_of(io + L.cvec(0)), Ry = cv_of(io + L.cvec(1)), Rz = cv_of(io + L.cvec(2));
    const CV u = cv_of(io + L.cvec(3)), v = cv_of(io + L.cvec(4)), w = cv_of(io + L.cvec(5)), P = cv_of(io + L.cvec(6));
    const SV s0{io + L.scratch(0), 1}, s1{io + L.scratch(1), 1}, s2{io + L.scratch(2), 1}, s3{io + L.scratch(3), 1};
    if (bulk) {
        double dPdx = 0.0, dPdz = 0.0;
        tau_solve_bulk(a.t, u, v, w, P, Rx, Ry, Rz, umean, wmean, s0, s1, s2, s3, dPdx, dPdz);
        io[L.dP()] = dPdx;
        io[L.dP() + 1] = dPdz;
    } else {
        tau_solve(a.t, u, v, w, P, Rx, Ry, Rz, s0, s1);
    }
}

// L2Norm2 of a Chebyshev series over [a,b], normalised (chebyshev.cpp:758-773)
__device__ double cheb_l2norm2(SV u, int N) {
    double sum = 0.0;
    for (int m = N - 1; m >= 0; --m) {
        double psum = 0.0;
        for (int n = m % 2; n < N; n += 2)
            psum += (u[n] * (1 - m * m - n * n)) / (double)((1 + m - n) * (1 - m + n) * (1 + m + n) * (1 - m - n));
        sum += u[m] * psum;
    }
    return sum;
}

// d^2/dy^2 by two passes of the derivative recurrence (chebyshev.cpp:672-708)
__device__ void cheb_diff2(SV u, SV d2, SV tmp, int N, double L) {
    cheb_diff(u, tmp, N, L);
    cheb_diff(tmp, d2, N, L);
}

// TauSolver::verify (tausolver.cpp:462-629): the sum of the L2 residuals of
// the three momentum equations, of the pressure equation and of the
// divergence, the wall values of u, v, v' (and u again, as the reference
// evaluates it in place of w), and the squared mean errors.
__global__ void k_ts_verify(TSParams a, double dPdx, double dPdz, double umean, double wmean, int mean_given) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const TSLayout Lo{a.Ny};
    const TauView& t = a.t;
    double* io = a.io;
    const int N = (int)a.Ny;
    const double L = t.b - t.a, lambda = t.vh.lambda, nu = t.nu, kappa2 = t.ph.lambda;
    const double kx = t.kx2pi, kz = t.kz2pi;
    const CV Rx = cv_of(io + Lo.cvec(0)), Ry = cv_of(io + Lo.cvec(1)), Rz = cv_of(io + Lo.cvec(2));
    const CV u = cv_of(io + Lo.cvec(3)), v = cv_of(io + Lo.cvec(4)), w = cv_of(io + Lo.cvec(5)), P = cv_of(io + Lo.cvec(6));
    const SV d2{io + Lo.scratch(0), 1}, tmp{io + Lo.scratch(1), 1}, lr{io + Lo.scratch(2), 1}, li{io + Lo.scratch(3), 1};
    if (!mean_given) {  // the 7-argument verify: umean, wmean from u, w, dPdx = dPdz = 0
        umean = cheb_mean(u.re, N);
        wmean = cheb_mean(w.re, N);
    }
    double error = 0.0;
    // u and w equations: lambda u - nu u'' + i k P (+ dP on the mean) against R
    auto momentum = [&](const CV& x, double k, double dP, const CV& R) {
        cheb_diff2(x.re, d2, tmp, N, L);
        for (int n = 0; n < N; ++n) lr[n] = lambda * x.re[n] - nu * d2[n] + (-(P.im[n] * k));
        cheb_diff2(x.im, d2, tmp, N, L);
        for (int n = 0; n < N; ++n) li[n] = lambda * x.im[n] - nu * d2[n] + P.re[n] * k;
        lr[0] += dP;
        for (int n = 0; n < N; ++n) {
            lr[n] -= R.re[n];
            li[n] -= R.im[n];
        }
        return sqrt(cheb_l2norm2(lr, N) + cheb_l2norm2(li, N));
    };
    error += momentum(u, kx, dPdx, Rx);
    {  // v equation: lambda v - nu v'' + P' against Ry
        const SV* parts[2] = {&lr, &li};
        for (int p = 0; p < 2; ++p) {
            const SV vv = p == 0 ? v.re : v.im, PP = p == 0 ? P.re : P.im, RR = p == 0 ? Ry.re : Ry.im;
            cheb_diff2(vv, d2, tmp, N, L);
            for (int n = 0; n < N; ++n) (*parts[p])[n] = lambda * vv[n] - nu * d2[n];
            cheb_diff(PP, tmp, N, L);
            for (int n = 0; n < N; ++n) {
                (*parts[p])[n] += tmp[n];
                (*parts[p])[n] -= RR[n];
            }
        }
        error += sqrt(cheb_l2norm2(lr, N) + cheb_l2norm2(li, N));
    }
    error += momentum(w, kz, dPdz, Rz);
    {  // pressure equation P'' - kappa^2 P against div R
        const SV* parts[2] = {&lr, &li};
        for (int p = 0; p < 2; ++p) {
            const SV PP = p == 0 ? P.re : P.im, RR = p == 0 ? Ry.re : Ry.im;
            cheb_diff2(PP, d2, tmp, N, L);
            for (int n = 0; n < N; ++n) (*parts[p])[n] = d2[n] - kappa2 * PP[n];
            cheb_diff(RR, tmp, N, L);
            for (int n = 0; n < N; ++n) {
                const double r = p == 0 ? tmp[n] - (kx * Rx.im[n] + kz * Rz.im[n]) : tmp[n] + (kx * Rx.re[n] + kz * Rz.re[n]);
                (*parts[p])[n] -= r;
            }
        }
        error += sqrt(cheb_l2norm2(lr, N) + cheb_l2norm2(li, N));
    }
    {  // divergence v' + i (kx u + kz w)
        cheb_diff(v.re, tmp, N, L);
        for (int n = 0; n < N; ++n) lr[n] = tmp[n] - (kx * u.im[n] + kz * w.im[n]);
        cheb_diff(v.im, tmp, N, L);
        for (int n = 0; n < N; ++n) li[n] = tmp[n] + (kx * u.re[n] + kz * w.re[n]);
        error += sqrt(cheb_l2norm2(lr, N) + cheb_l2norm2(li, N));
    }
    auto cabs = [](double re, double im) { return sqrt(re * re + im * im); };
    const double uar = cheb_eval_a(u.re, N), uai = cheb_eval_a(u.im, N);
    const double ubr = cheb_eval_b(u.re, N), ubi = cheb_eval_b(u.im, N);
    error += cabs(uar, uai) + cabs(ubr, ubi);
    error += cabs(cheb_eval_a(v.re, N), cheb_eval_a(v.im, N)) + cabs(cheb_eval_b(v.re, N), cheb_eval_b(v.im, N));
    double vyar, vybr, vyai, vybi;
    cheb_diff_ends(v.re, tmp, N, L, vybr, vyar);
    cheb_diff_ends(v.im, tmp, N, L, vybi, vyai);
    error += cabs(vyar, vyai) + cabs(vybr, vybi);
    error += cabs(uar, uai) + cabs(ubr, ubi);  // tausolver.cpp:599-600 evaluates u here
    const double eu = cheb_mean(u.re, N) - umean, ew = cheb_mean(w.re, N) - wmean;
    error += eu * eu + ew * ew;
    io[Lo.err()] = error;
}

}  // namespace

struct cfg_tausolver {
    int kx = 0, kz = 0, Ny = 0, taucorr = 1;
    double Lx = 0, Lz = 0, a = -1, b = 1, lambda = 0, nu = 0;
    DBuf tau;  // TAU_NA Ny-vectors and 6 scalars
    DBuf io;
    hipStream_t st = nullptr;

    TauView view() {
        const long N = Ny;
        double* cm = tau.p;
        auto A = [&](int k) { return SV{cm + (long)k * N, 1}; };
        const double kappa2 = 4 * (kPi * kPi) * ((kx / Lx) * (kx / Lx) + (kz / Lz) * (kz / Lz));
        TauView t;
        t.ph = HelmholtzView{A(0), A(1), A(2), kappa2, Ny};
        t.vh = HelmholtzView{A(3), A(4), A(5), lambda, Ny};
        t.Pp = A(6);
        t.Pm = A(7);
        t.vp = A(8);
        t.vm = A(9);
        t.P0 = A(10);
        t.v0 = A(11);
        t.scal = cm + (long)TAU_NA * N;
        t.sstride = 1;
        t.nu = nu;
        t.a = a;
        t.b = b;
        t.kx2pi = 2 * kPi * kx / Lx;
        t.kz2pi = 2 * kPi * kz / Lz;
        t.zero_mode = (kx == 0 && kz == 0);
        t.taucorrection = taucorr != 0;
        return t;
    }
    TSParams params() { return TSParams{view(), io.p, Ny}; }

    void put(const double* const* src, int first, int count) {  // complex Ny-vectors into io slots
        const TSLayout L{Ny};
        for (int k = 0; k < count; ++k)
            if (src[k])
                HIPCHK(hipMemcpyAsync(io.p + L.cvec(first + k), src[k], 2 * Ny * sizeof(double), hipMemcpyHostToDevice,
                                      st));
    }
    void get(double* const* dst, int first, int count) {
        const TSLayout L{Ny};
        for (int k = 0; k < count; ++k)
            if (dst[k])
                HIPCHK(hipMemcpyAsync(dst[k], io.p + L.cvec(first + k), 2 * Ny * sizeof(double), hipMemcpyDeviceToHost,
                                      st));
    }
};

static void check_vectors(std::initializer_list<const double*> v) {
    for (const double* p : v)
        if (!p) throw Error(CFG_ERR_ARG, "null argument");
}

extern "C" {

int cfg_tausolver_create(int kx, int kz, double Lx, double Lz, double a, double b, double lambda, double nu, int Ny,
                         int taucorrection, cfg_tausolver** out) {
    return guarded([&] {
        if (!out) throw Error(CFG_ERR_ARG, "null argument");
        if (Ny < 5 || Ny % 2 == 0) throw Error(CFG_ERR_ARG, "Ny must be odd and >= 5");
        if (!(Lx > 0) || !(Lz > 0) || !(b > a) || !(nu > 0)) throw Error(CFG_ERR_ARG, "bad box or nu");
        auto s = std::make_unique<cfg_tausolver>();
        s->kx = kx;
        s->kz = kz;
        s->Ny = Ny;
        s->taucorr = taucorrection != 0;
        s->Lx = Lx;
        s->Lz = Lz;
        s->a = a;
        s->b = b;
        s->lambda = lambda;
        s->nu = nu;
        HIPCHK(hipStreamCreate(&s->st));
        s->tau.alloc((size_t)TAU_NA * Ny + 6);
        const TSLayout L{Ny};
        s->io.alloc((size_t)L.total());
        HIPCHK(hipMemsetAsync(s->io.p, 0, (size_t)L.total() * sizeof(double), s->st));
        k_ts_setup<<<1, 64, 0, s->st>>>(s->params());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->st));
        *out = s.release();
    });
}

int cfg_tausolver_destroy(cfg_tausolver* s) {
    return guarded([&] {
        if (!s) return;
        (void)hipStreamSynchronize(s->st);
        hipStream_t st = s->st;
        delete s;
        (void)hipStreamDestroy(st);
    });
}

int cfg_tausolver_solve(cfg_tausolver* s, const double* Rx, const double* Ry, const double* Rz, double* u, double* v,
                        double* w, double* P) {
    return guarded([&] {
        if (!s) throw Error(CFG_ERR_ARG, "null solver");
        check_vectors({Rx, Ry, Rz, u, v, w, P});
        const double* in[3] = {Rx, Ry, Rz};
        s->put(in, 0, 3);
        k_ts_solve<<<1, 64, 0, s->st>>>(s->params(), 0, 0.0, 0.0);
        HIPCHK(hipGetLastError());
        double* o[4] = {u, v, w, P};
        s->get(o, 3, 4);
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

int cfg_tausolver_solve_bulk(cfg_tausolver* s, const double* Rx, const double* Ry, const double* Rz, double umean,
                             double wmean, double* u, double* v, double* w, double* P, double* dPdx, double* dPdz) {
    return guarded([&] {
        if (!s) throw Error(CFG_ERR_ARG, "null solver");
        check_vectors({Rx, Ry, Rz, u, v, w, P});
        if (s->kx != 0 || s->kz != 0)
            throw Error(CFG_ERR_ARG, "the bulk-velocity solve is for kx = kz = 0 only (tausolver.cpp:407-410)");
        const double* in[3] = {Rx, Ry, Rz};
        s->put(in, 0, 3);
        k_ts_solve<<<1, 64, 0, s->st>>>(s->params(), 1, umean, wmean);
        HIPCHK(hipGetLastError());
        double* o[4] = {u, v, w, P};
        s->get(o, 3, 4);
        double dP[2];
        const TSLayout L{s->Ny};
        HIPCHK(hipMemcpyAsync(dP, s->io.p + L.dP(), 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        if (dPdx) *dPdx = dP[0];
        if (dPdz) *dPdz = dP[1];
    });
}

static int verify_impl(cfg_tausolver* s, const double* u, const double* v, const double* w, const double* P,
                       const double* Rx, const double* Ry, const double* Rz, double dPdx, double dPdz, double umean,
                       double wmean, int mean_given, double* error) {
    return guarded([&] {
        if (!s || !error) throw Error(CFG_ERR_ARG, "null argument");
        check_vectors({Rx, Ry, Rz, u, v, w, P});
        const double* in[7] = {Rx, Ry, Rz, u, v, w, P};
        s->put(in, 0, 7);
        k_ts_verify<<<1, 64, 0, s->st>>>(s->params(), dPdx, dPdz, umean, wmean, mean_given);
        HIPCHK(hipGetLastError());
        const TSLayout L{s->Ny};
        HIPCHK(hipMemcpyAsync(error, s->io.p + L.err(), sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

int cfg_tausolver_verify(cfg_tausolver* s, const double* u, const double* v, const double* w, const double* P,
                         const double* Rx, const double* Ry, const double* Rz, double* error) {
    return verify_impl(s, u, v, w, P, Rx, Ry, Rz, 0.0, 0.0, 0.0, 0.0, 0, error);
}

int cfg_tausolver_verify_bulk(cfg_tausolver* s, const double* u, const double* v, const double* w, const double* P,
                              double dPdx, double dPdz, const double* Rx, const double* Ry, const double* Rz,
                              double umean, double wmean, double* error) {
    return verify_impl(s, u, v, w, P, Rx, Ry, Rz, dPdx, dPdz, umean, wmean, 1, error);
}

}  // extern "C"
