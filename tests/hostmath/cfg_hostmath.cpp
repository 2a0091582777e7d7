// Host build of the product's per-mode math (channelflow_amd/csrc/cfg_math.hpp,
// the same functions the tau-solve kernels run), exposed to the CPU test suite.
#include <vector>

#include "../../channelflow_amd/csrc/cfg_math.hpp"

using namespace cfg;

extern "C" {

// One mode: TauSolver construction + solve, profiles contiguous (stride 1).
// R: Rx,Ry,Rz interleaved complex (3*Ny); out: u,v,w,P interleaved (4*Ny).
int hm_tausolve(int kx, int kz, double Lx, double Lz, double a, double b, double lambda_t, double nu, int Ny,
                int taucorr, const double* R, double* out) {
    const double PI = 3.14159265358979323846;
    const double pi2 = PI * PI;
    const double kappa2 = 4 * pi2 * ((kx / Lx) * (kx / Lx) + (kz / Lz) * (kz / Lz));
    const double lambda = lambda_t + 4.0 * pi2 * nu * ((kx / Lx) * (kx / Lx) + (kz / Lz) * (kz / Lz));
    std::vector<std::vector<double>> A(12, std::vector<double>(Ny, 0.0));
    std::vector<double> scal(6, 0.0), s0(Ny), s1(Ny);
    auto sv = [](std::vector<double>& v) { return SV{v.data(), 1}; };
    TauView t;
    t.ph = HelmholtzView{sv(A[0]), sv(A[1]), sv(A[2]), kappa2, Ny};
    t.vh = HelmholtzView{sv(A[3]), sv(A[4]), sv(A[5]), lambda, Ny};
    t.Pp = sv(A[6]);
    t.Pm = sv(A[7]);
    t.vp = sv(A[8]);
    t.vm = sv(A[9]);
    t.P0 = sv(A[10]);
    t.v0 = sv(A[11]);
    t.scal = scal.data();
    t.sstride = 1;
    t.nu = nu;
    t.a = a;
    t.b = b;
    t.kx2pi = 2 * PI * kx / Lx;
    t.kz2pi = 2 * PI * kz / Lz;
    t.zero_mode = (kx == 0 && kz == 0);
    t.taucorrection = taucorr != 0;
    tau_setup(t, sv(s0), sv(s1));
    std::vector<double> in(R, R + 6 * Ny);
    std::vector<double> o(8 * Ny, 0.0);
    auto cv = [&](std::vector<double>& base, int comp) {
        return CV{SV{base.data() + 2 * comp * Ny, 2}, SV{base.data() + 2 * comp * Ny + 1, 2}};
    };
    tau_solve(t, cv(o, 0), cv(o, 1), cv(o, 2), cv(o, 3), cv(in, 0), cv(in, 1), cv(in, 2), sv(s0), sv(s1));
    for (int i = 0; i < 8 * Ny; ++i) out[i] = o[i];
    return 0;
}

// HelmholtzSolver::solve via the packed factors.
int hm_helmholtz(int Ny, double a, double b, double lambda, double nu, const double* f, double ua, double ub,
                 double* u) {
    std::vector<double> band(Ny), up(Ny), invd(Ny), ff(f, f + Ny), uu(Ny);
    HelmholtzView h{SV{band.data(), 1}, SV{up.data(), 1}, SV{invd.data(), 1}, lambda, Ny};
    helm_factor(h, nu, a, b);
    helm_solve(h, SV{uu.data(), 1}, SV{ff.data(), 1}, ua, ub);
    for (int i = 0; i < Ny; ++i) u[i] = uu[i];
    return 0;
}

}  // extern "C"
