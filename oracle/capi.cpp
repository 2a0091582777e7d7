// extern "C" surface of the oracle, for ctypes in tests/ and bench.py's
// cpu_baseline leg. TEST INFRASTRUCTURE ONLY -- never linked by the product.
//
// Spectral buffers use channelflow's serial complex layout, interleaved re/im:
//   index(mx,my,mz,i) = mz + (Nz/2+1)*(mx + Nx*(my + Ny*i))     (flowfield.h:400)
#include <cmath>
#include <cstring>
#include <exception>
#include <string>

#include "cpu_ref.hpp"

using namespace cfref;

static thread_local std::string g_err;

template <class R, class F>
static R guard(R fail, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return fail;
    }
}

extern "C" {

struct cpuref_flags {  // field order shared with the product's cfg_flags (include/cfgpu.h)
    int baseflow, constraint, timestepping, initstepping, nonlinearity, dealiasing, taucorrection, reserved;
    double nu, Vsuck, rotation, dPdx, dPdz, Ubulk, Wbulk, ulowerwall, uupperwall, wlowerwall, wupperwall, t0, dt;
};

const char* cpuref_last_error() { return g_err.c_str(); }

static DNSFlags toflags(const cpuref_flags* f) {
    DNSFlags d;
    d.baseflow = (BaseFlow)f->baseflow;
    d.constraint = (MeanConstraint)f->constraint;
    d.timestepping = (TimeStepMethod)f->timestepping;
    d.initstepping = (TimeStepMethod)f->initstepping;
    d.nonlinearity = (NonlinearMethod)f->nonlinearity;
    d.dealiasing = (Dealiasing)f->dealiasing;
    d.taucorrection = f->taucorrection != 0;
    d.nu = f->nu; d.Vsuck = f->Vsuck; d.rotation = f->rotation; d.dPdx = f->dPdx; d.dPdz = f->dPdz;
    d.Ubulk = f->Ubulk; d.Wbulk = f->Wbulk; d.ulowerwall = f->ulowerwall; d.uupperwall = f->uupperwall;
    d.wlowerwall = f->wlowerwall; d.wupperwall = f->wupperwall; d.t0 = f->t0; d.dt = f->dt;
    return d;
}

static void load(FlowField& f, const double* spec) {
    std::memcpy(f.rdata(), spec, sizeof(double) * f.Nloc());
    f.setState(Spectral, Spectral);
}
static void store(const FlowField& f0, double* spec) {
    FlowField f(f0);
    f.makeSpectral();
    std::memcpy(spec, f.rdata(), sizeof(double) * f.Nloc());
}

struct Handle {
    std::vector<FlowField> fields;
    std::unique_ptr<DNS> dns;
};


void* cpuref_dns_create(int Nx, int Ny, int Nz, double Lx, double Lz, double a, double b, const cpuref_flags* fl,
                        const double* u_spec, const double* q_spec, int padded, const double* Ubase,
                        const double* Wbase) {
    return guard<void*>(nullptr, [&]() -> void* {
            auto* h = new Handle;
            h->fields = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(h->fields[0], u_spec);
            if (q_spec) load(h->fields[1], q_spec);
            h->fields[0].setPadded(padded != 0);
            h->fields[1].setPadded(padded != 0);
            DNSFlags flags = toflags(fl);
            if (Ubase) {
                ChebyCoeff U(Ny, a, b), W(Ny, a, b);
                for (int i = 0; i < Ny; ++i) { U[i] = Ubase[i]; W[i] = Wbase ? Wbase[i] : 0.0; }
                h->dns = std::make_unique<DNS>(h->fields, std::vector<ChebyCoeff>{U, W}, flags);
            } else {
                h->dns = std::make_unique<DNS>(h->fields, flags);
            }
            return (void*)h;
        });
}

int cpuref_dns_advance(void* hp, int nsteps) {
    return guard<int>(-1, [&]() -> int { static_cast<Handle*>(hp)->dns->advance(static_cast<Handle*>(hp)->fields, nsteps); return 0; });
}

int cpuref_dns_reset_dt(void* hp, double dt) {
    return guard<int>(-1, [&]() -> int { static_cast<Handle*>(hp)->dns->reset_dt(dt); return 0; });
}

void cpuref_dns_get(void* hp, double* u_spec, double* q_spec) {
    auto* h = static_cast<Handle*>(hp);
    if (u_spec) store(h->fields[0], u_spec);
    if (q_spec) store(h->fields[1], q_spec);
}

// The caller's fields between advance() calls (DNS::advance copies them into
// its history, dnsalgo.cpp:207): as tests/dnsParabolaTest.cpp:177-178 sets u
// after the DNS was built from zero fields.
void cpuref_dns_set(void* hp, const double* u_spec, const double* q_spec) {
    auto* h = static_cast<Handle*>(hp);
    const bool pu = h->fields[0].padded(), pq = h->fields[1].padded();
    if (u_spec) load(h->fields[0], u_spec);
    if (q_spec) load(h->fields[1], q_spec);
    h->fields[0].setPadded(pu);
    h->fields[1].setPadded(pq);
}

double cpuref_dns_time(void* hp) { return static_cast<Handle*>(hp)->dns->time(); }
double cpuref_dns_dpdx(void* hp) { return static_cast<Handle*>(hp)->dns->dPdx(); }
double cpuref_dns_ubulk(void* hp) { return static_cast<Handle*>(hp)->dns->Ubulk(); }
double cpuref_dns_cfl(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    return h->dns->CFL(h->fields[0]);
}
void cpuref_dns_base(void* hp, double* U, double* W) {
    auto* h = static_cast<Handle*>(hp);
    const ChebyCoeff& Ub = h->dns->Ubase();
    const ChebyCoeff& Wb = h->dns->Wbase();
    for (int i = 0; i < Ub.N(); ++i) { U[i] = Ub[i]; W[i] = Wb[i]; }
}
void cpuref_dns_destroy(void* hp) { delete static_cast<Handle*>(hp); }

// PressureSolver(u, dns.Ubase(), dns.Wbase(), nu, Vsuck, nonlinearity).solve(p, u),
// as simulateflow.cpp:100-105 calls it: the base flow is the one a DNS with
// these flags builds.
int cpuref_pressure(int Nx, int Ny, int Nz, double Lx, double Lz, double a, double b, const cpuref_flags* fl,
                    const double* u_spec, double* p_spec) {
    return guard<int>(-1, [&]() -> int {
            std::vector<FlowField> in = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(in[0], u_spec);
            const DNSFlags f = toflags(fl);
            NSE nse(in, f);
            FlowField p;
            pressure_solve(in[0], nse.Ubase(), nse.Wbase(), f.nu, f.Vsuck, f.nonlinearity, p);
            store(p, p_spec);
            return 0;
        });
}

// One NSE::nonlinear call (nse.cpp:383-391): base flow added, NL evaluated,
// aliased modes of f zeroed. u is round-tripped in place as in the reference.
int cpuref_nonlinear(int Nx, int Ny, int Nz, double Lx, double Lz, double a, double b, const cpuref_flags* fl,
                     double* u_spec, double* f_spec) {
    return guard<int>(-1, [&]() -> int {
            std::vector<FlowField> in = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(in[0], u_spec);
            std::vector<FlowField> out = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b)};
            NSE nse(in, toflags(fl));
            nse.nonlinear(in, out);
            store(in[0], u_spec);
            store(out[0], f_spec);
            return 0;
        });
}

// One NSE::linear call (nse.cpp:393-477); out is pre-filled with u as
// CNABstyleDNS::advance does (lt is a copy of the RHS template).
int cpuref_linear(int Nx, int Ny, int Nz, double Lx, double Lz, double a, double b, const cpuref_flags* fl,
                  const double* u_spec, const double* q_spec, double* l_spec) {
    return guard<int>(-1, [&]() -> int {
            std::vector<FlowField> in = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(in[0], u_spec);
            load(in[1], q_spec);
            std::vector<FlowField> out = {in[0]};
            NSE nse(in, toflags(fl));
            nse.linear(in, out);
            store(out[0], l_spec);
            return 0;
        });
}

// One NSE::solve call (nse.cpp:479-575) with TauSolvers built for lambda_t.
// u_out/q_out are pre-filled by the caller (non-live modes are left as given).
int cpuref_solve(int Nx, int Ny, int Nz, double Lx, double Lz, double a, double b, const cpuref_flags* fl,
                 double lambda_t, const double* u0_spec, const double* rhs_spec, double* u_out, double* q_out,
                 double* dPdx_out) {
    return guard<int>(-1, [&]() -> int {
            std::vector<FlowField> init = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(init[0], u0_spec);
            NSE nse(init, toflags(fl));
            nse.reset_lambda({lambda_t});
            std::vector<FlowField> out = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b), FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b)};
            load(out[0], u_out);
            load(out[1], q_out);
            std::vector<FlowField> rhs = {FlowField(Nx, Ny, Nz, 3, Lx, Lz, a, b)};
            load(rhs[0], rhs_spec);
            nse.solve(out, rhs, 0);
            store(out[0], u_out);
            store(out[1], q_out);
            if (dPdx_out) *dPdx_out = nse.dPdx();
            return 0;
        });
}

// TauSolver for one mode (tausolver.cpp:81-402). R holds Rx,Ry,Rz as
// interleaved complex Ny-vectors; out receives u,v,w,P likewise.
int cpuref_tausolve(int kx, int kz, double Lx, double Lz, double a, double b, double lambda, double nu, int Ny,
                    int taucorrection, const double* R, double* out) {
    return guard<int>(-1, [&]() -> int {
            TauSolver ts(kx, kz, Lx, Lz, a, b, lambda, nu, Ny, taucorrection != 0);
            ComplexChebyCoeff Rx(Ny, a, b), Ry(Ny, a, b), Rz(Ny, a, b), u(Ny, a, b), v(Ny, a, b), w(Ny, a, b),
                P(Ny, a, b);
            for (int n = 0; n < Ny; ++n) {
                Rx.set(n, Complex(R[2 * n], R[2 * n + 1]));
                Ry.set(n, Complex(R[2 * Ny + 2 * n], R[2 * Ny + 2 * n + 1]));
                Rz.set(n, Complex(R[4 * Ny + 2 * n], R[4 * Ny + 2 * n + 1]));
            }
            ts.solve(u, v, w, P, Rx, Ry, Rz);
            ComplexChebyCoeff* o[4] = {&u, &v, &w, &P};
            for (int c = 0; c < 4; ++c)
                for (int n = 0; n < Ny; ++n) {
                    out[2 * (c * Ny + n)] = o[c]->re[n];
                    out[2 * (c * Ny + n) + 1] = o[c]->im[n];
                }    return 0;
    });
}

// TauSolver's bulk-velocity solve for kx = kz = 0 (tausolver.cpp:404-450): as
// cpuref_tausolve, plus dP[0] = dPdx and dP[1] = dPdz.
int cpuref_tausolve_bulk(double Lx, double Lz, double a, double b, double lambda, double nu, int Ny,
                         int taucorrection, const double* R, double umean, double wmean, double* out, double* dP) {
    return guard<int>(-1, [&]() -> int {
        TauSolver ts(0, 0, Lx, Lz, a, b, lambda, nu, Ny, taucorrection != 0);
        ComplexChebyCoeff Rx(Ny, a, b), Ry(Ny, a, b), Rz(Ny, a, b), u(Ny, a, b), v(Ny, a, b), w(Ny, a, b),
            P(Ny, a, b);
        for (int n = 0; n < Ny; ++n) {
            Rx.set(n, Complex(R[2 * n], R[2 * n + 1]));
            Ry.set(n, Complex(R[2 * Ny + 2 * n], R[2 * Ny + 2 * n + 1]));
            Rz.set(n, Complex(R[4 * Ny + 2 * n], R[4 * Ny + 2 * n + 1]));
        }
        ts.solve(u, v, w, P, dP[0], dP[1], Rx, Ry, Rz, umean, wmean);
        ComplexChebyCoeff* o[4] = {&u, &v, &w, &P};
        for (int c = 0; c < 4; ++c)
            for (int n = 0; n < Ny; ++n) {
                out[2 * (c * Ny + n)] = o[c]->re[n];
                out[2 * (c * Ny + n) + 1] = o[c]->im[n];
            }
        return 0;
    });
}

// HelmholtzSolver (helmholtz.cpp:18-95): nu u'' - lambda u = f, u(a)=ua, u(b)=ub.
int cpuref_helmholtz(int Ny, double a, double b, double lambda, double nu, const double* f, double ua, double ub,
                     double* u) {
    return guard<int>(-1, [&]() -> int {
            HelmholtzSolver h(Ny, a, b, lambda, nu);
            ChebyCoeff F(Ny, a, b), U(Ny, a, b);
            for (int i = 0; i < Ny; ++i) F[i] = f[i];
            h.solve(U, F, ua, ub);
            for (int i = 0; i < Ny; ++i) u[i] = U[i];
            return 0;
        });
}

// Mean-constrained Helmholtz (helmholtz.cpp:158-213).
int cpuref_helmholtz_mean(int Ny, double a, double b, double lambda, double nu, const double* f, double umean,
                          double ua, double ub, double* u, double* mu) {
    return guard<int>(-1, [&]() -> int {
            HelmholtzSolver h(Ny, a, b, lambda, nu);
            ChebyCoeff F(Ny, a, b), U(Ny, a, b);
            for (int i = 0; i < Ny; ++i) F[i] = f[i];
            h.solve(U, *mu, F, umean, ua, ub);
            for (int i = 0; i < Ny; ++i) u[i] = U[i];
            return 0;
        });
}

// BandedTridiag UL solve (bandedtridiag.cpp:212-277) on a dense description:
// band[M], lo[M], diag[M], up[M]; b overwritten with A^{-1} b.
int cpuref_banded_solve(int M, const double* band, const double* lo, const double* dg, const double* up, double* b) {
    return guard<int>(-1, [&]() -> int {
            BandedTridiag A(M);
            for (int i = 0; i < M; ++i) {
                A.band(i) = band[i];
                A.lodiag(i) = lo[i];
                A.diag(i) = dg[i];
                A.updiag(i) = up[i];
            }
            A.diag(0) = band[0];
            A.ULdecomp();
            A.ULsolveStrided(b, 0, 1);
            return 0;
        });
}

// Chebyshev primitives (chebyshev.cpp:262-303, 405-512, 672-708, 758).
void cpuref_cheb_diff(int N, double a, double b, const double* u, double* du) {
    ChebyCoeff U(N, a, b), D(N, a, b);
    for (int i = 0; i < N; ++i) U[i] = u[i];
    diff(U, D);
    for (int i = 0; i < N; ++i) du[i] = D[i];
}
void cpuref_cheb_spectral(int N, double* x) {  // physical values at chebypoints -> coefficients
    ChebyCoeff U(N, -1, 1, Physical);
    for (int i = 0; i < N; ++i) U[i] = x[i];
    U.makeSpectral();
    for (int i = 0; i < N; ++i) x[i] = U[i];
}
void cpuref_cheb_physical(int N, double* x) {
    ChebyCoeff U(N, -1, 1, Spectral);
    for (int i = 0; i < N; ++i) U[i] = x[i];
    U.makePhysical();
    for (int i = 0; i < N; ++i) x[i] = U[i];
}
double cpuref_cheb_mean(int N, const double* u) {
    ChebyCoeff U(N, -1, 1);
    for (int i = 0; i < N; ++i) U[i] = u[i];
    return U.mean();
}
double cpuref_cheb_l2norm2(int N, double a, double b, const double* u) {
    ChebyCoeff U(N, a, b);
    for (int i = 0; i < N; ++i) U[i] = u[i];
    return L2Norm2(U, true);
}

// Whole-field transforms: physical real layout <-> spectral complex layout,
// same buffer (flowfield.cpp:1850-1997).
void cpuref_field_make_spectral(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b, double* d) {
    FlowField f(Nx, Ny, Nz, Nd, Lx, Lz, a, b, Physical, Physical);
    std::memcpy(f.rdata(), d, sizeof(double) * f.Nloc());
    f.makeSpectral();
    std::memcpy(d, f.rdata(), sizeof(double) * f.Nloc());
}
void cpuref_field_make_physical(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b, double* d) {
    FlowField f(Nx, Ny, Nz, Nd, Lx, Lz, a, b, Spectral, Spectral);
    std::memcpy(f.rdata(), d, sizeof(double) * f.Nloc());
    f.makePhysical();
    std::memcpy(d, f.rdata(), sizeof(double) * f.Nloc());
}

// L2Norm2 / L2Dist2 over a spectral field (diffops.cpp:353-487).
double cpuref_l2norm2(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b, const double* u,
                      int padded) {
    FlowField f(Nx, Ny, Nz, Nd, Lx, Lz, a, b);
    std::memcpy(f.rdata(), u, sizeof(double) * f.Nloc());
    f.setPadded(padded != 0);
    return L2Norm2(f, true);
}
double cpuref_l2dist2(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b, const double* u,
                      const double* v, int padded) {
    FlowField f(Nx, Ny, Nz, Nd, Lx, Lz, a, b), g(Nx, Ny, Nz, Nd, Lx, Lz, a, b);
    std::memcpy(f.rdata(), u, sizeof(double) * f.Nloc());
    std::memcpy(g.rdata(), v, sizeof(double) * g.Nloc());
    f.setPadded(padded != 0);
    g.setPadded(padded != 0);
    return L2Dist2(f, g, true);
}

// De-padded physical file data (component, z, y, x) on the 2(Nx/3) x Ny x
// 2(Nz/3) grid -> padded spectral field: flowfield.cpp:3605-3826 (readNetCDF,
// serial branch) + addPaddedModes (:2991-3215) + makeSpectral_y.
int cpuref_field_from_io(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b, const double* io,
                         double* out_spec) {
    return guard<int>(-1, [&]() -> int {
            const int Nxio = 2 * (Nx / 3), Nzio = 2 * (Nz / 3);
            FlowField g(Nxio, Ny, Nzio, Nd, Lx, Lz, a, b, Physical, Physical);
            for (int i = 0; i < Nd; ++i)
                for (int nz = 0; nz < Nzio; ++nz)
                    for (int ny = 0; ny < Ny; ++ny)
                        for (int nx = 0; nx < Nxio; ++nx)
                            g.rv(nx, ny, nz, i) = io[nx + Nxio * (ny + Ny * (nz + Nzio * i))];
            g.makeSpectral_xz();
            FlowField f(Nx, Ny, Nz, Nd, Lx, Lz, a, b, Spectral, Physical);
            const int Mzio = Nzio / 2 + 1;
            for (int i = 0; i < Nd; ++i)
                for (int ny = 0; ny < Ny; ++ny)
                    for (int mz = 0; mz < Mzio && mz < f.Mz(); ++mz) {
                        for (int mx = 0; mx < Nxio / 2 + 1; ++mx) f.cmplx(mx, ny, mz, i) = g.cmplx(mx, ny, mz, i);
                        for (int mx = Nxio / 2 + 1; mx < Nxio; ++mx)
                            f.cmplx(mx + (Nx - Nxio), ny, mz, i) = g.cmplx(mx, ny, mz, i);
                    }
            f.makeSpectral_y();
            std::memcpy(out_spec, f.rdata(), sizeof(double) * f.Nloc());
            return 0;
        });
}

int cpuref_laminar_profile(double nu, int constraint, double dPdx, double Ubulk, double Vsuck, double a, double b,
                           double ua, double ub, int Ny, double* out) {
    return guard<int>(-1, [&]() -> int {
            ChebyCoeff u = laminarProfile(nu, (MeanConstraint)constraint, dPdx, Ubulk, Vsuck, a, b, ua, ub, Ny);
            for (int i = 0; i < Ny; ++i) out[i] = u[i];
            return 0;
        });
}

}  // extern "C"
