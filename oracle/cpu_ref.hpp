// cpu_ref -- FP64 CPU restatement of channelflow's DNS time step, used ONLY as
// the parity oracle and the timed CPU baseline.
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only to check results against. The
// product (channelflow_amd, libcfgpu.so) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the channelflow 2.0.2 source tree). The FFTs are this file's own mixed-radix
// transforms (FFTW is absent here), so results agree with the reference to
// rounding, not bit-for-bit.
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace cfref {

using Real = double;
using Complex = std::complex<double>;
constexpr Real pi = 3.14159265358979323846;  // cfbasics/mathdefs.h:84

enum FieldState { Physical = 0, Spectral = 1 };

// ---- enums, numbered as channelflow/dnsflags.h:26-38 ----------------------
enum BaseFlow { ZeroBase, LinearBase, ParabolicBase, LaminarBase, SuctionBase, ArbitraryBase };
enum MeanConstraint { PressureGradient, BulkVelocity };
enum TimeStepMethod { CNFE1, CNAB2, CNRK2, SMRK2, SBDF1, SBDF2, SBDF3, SBDF4 };
enum NonlinearMethod { Rotational, Convection, Divergence, SkewSymmetric, Alternating, Alternating_, LinearAboutProfile };
enum Dealiasing { NoDealiasing, DealiasXZ, DealiasY, DealiasXYZ };

// DNSFlags fields on the DNS::advance path (dnsflags.h:84-140; defaults :84-92)
struct DNSFlags {
    BaseFlow baseflow = LaminarBase;
    MeanConstraint constraint = PressureGradient;
    TimeStepMethod timestepping = SBDF3;
    TimeStepMethod initstepping = SMRK2;
    NonlinearMethod nonlinearity = Rotational;
    Dealiasing dealiasing = DealiasXZ;
    bool taucorrection = true;
    Real nu = 0.0025, Vsuck = 0.0, rotation = 0.0, dPdx = 0.0, dPdz = 0.0, Ubulk = 0.0, Wbulk = 0.0;
    Real ulowerwall = 0.0, uupperwall = 0.0, wlowerwall = 0.0, wupperwall = 0.0;
    Real t0 = 0.0, dt = 0.03125;
    bool dealias_xz() const { return dealiasing == DealiasXZ || dealiasing == DealiasXYZ; }  // dnsflags.cpp:269
    bool dealias_y() const { return dealiasing == DealiasY || dealiasing == DealiasXYZ; }    // dnsflags.cpp:271
};

// ---- FFTs (own implementation; stand-in for FFTW's DFT and REDFT00) -------
class CFFT {  // unnormalised complex DFT, sign -1 forward, +1 backward
   public:
    explicit CFFT(int n);
    void exec(Complex* x, int sign) const;  // in place, contiguous
    int n() const { return n_; }

   private:
    void rec(Complex* out, const Complex* in, int nsub, int stride, int fi, int sign) const;
    int n_;
    std::vector<int> fac_;
    std::vector<Complex> w_;  // exp(-2 pi i k / n)
    mutable std::vector<Complex> scratch_;
};

class RFFT {  // real <-> half-complex, n even, unnormalised (FFTW r2c / c2r conventions)
   public:
    explicit RFFT(int n);
    void r2c(const Real* in, Complex* out) const;  // out[0..n/2]
    void c2r(const Complex* in, Real* out) const;  // imag of in[0], in[n/2] ignored
   private:
    int n_;
    CFFT half_;
    std::vector<Complex> tw_;  // exp(-2 pi i k / n), k <= n/2
    mutable std::vector<Complex> z_;
};

class DCT1 {  // FFTW_REDFT00: Y_k = x0 + (-1)^k x_{N-1} + 2 sum_{j=1}^{N-2} x_j cos(pi j k/(N-1))
   public:
    explicit DCT1(int N);
    void exec(Real* x) const;

   private:
    int N_;
    RFFT r_;
    mutable std::vector<Real> e_;
    mutable std::vector<Complex> E_;
};

// ---- Chebyshev coefficients (chebyshev.h/.cpp) ------------------------------
struct ChebyCoeff {
    std::vector<Real> d;
    Real a = -1, b = 1;
    FieldState state = Spectral;
    ChebyCoeff() = default;
    ChebyCoeff(int N, Real a_, Real b_, FieldState s = Spectral) : d(N, 0.0), a(a_), b(b_), state(s) {}
    int N() const { return (int)d.size(); }
    Real L() const { return b - a; }
    Real& operator[](int i) { return d[i]; }
    const Real& operator[](int i) const { return d[i]; }
    void setToZero() { std::fill(d.begin(), d.end(), 0.0); }
    Real eval_a() const;  // chebyshev.cpp:418
    Real eval_b() const;  // chebyshev.cpp:405
    Real mean() const;    // chebyshev.cpp:505
    void makeSpectral();  // chebyshev.cpp:262
    void makePhysical();  // chebyshev.cpp:286
};
void diff(const ChebyCoeff& u, ChebyCoeff& du);               // chebyshev.cpp:672
void diff2(const ChebyCoeff& u, ChebyCoeff& d2u);             // chebyshev.cpp:699
Real L2Norm2(const ChebyCoeff& u, bool normalize = true);     // chebyshev.cpp:758
std::vector<Real> chebypoints(int N, Real a, Real b);         // chebyshev.cpp:746

struct ComplexChebyCoeff {
    ChebyCoeff re, im;
    ComplexChebyCoeff() = default;
    ComplexChebyCoeff(int N, Real a, Real b) : re(N, a, b), im(N, a, b) {}
    int N() const { return re.N(); }
    Complex get(int i) const { return Complex(re[i], im[i]); }
    void set(int i, Complex c) { re[i] = c.real(); im[i] = c.imag(); }
};
void diff(const ComplexChebyCoeff& u, ComplexChebyCoeff& du);

// ---- banded tridiagonal (bandedtridiag.h/.cpp) ------------------------------
class BandedTridiag {
   public:
    BandedTridiag() = default;
    explicit BandedTridiag(int M);
    Real& band(int j) { return band_[j]; }
    Real& diag(int i) { return diag_[i]; }
    Real& updiag(int i) { return up_[i]; }
    Real& lodiag(int i) { return lo_[i]; }
    void ULdecomp();                                                           // bandedtridiag.cpp:212
    void ULsolveStrided(Real* b, int offset, int stride) const;               // :232
    void multiplyStrided(const Real* x, Real* b, int offset, int stride) const;  // :315
    int M() const { return M_; }
    const std::vector<Real>& bandv() const { return band_; }
    const std::vector<Real>& diagv() const { return diag_; }
    const std::vector<Real>& upv() const { return up_; }
    const std::vector<Real>& lov() const { return lo_; }
    const std::vector<Real>& invdiagv() const { return invdiag_; }

   private:
    int M_ = 0;
    std::vector<Real> band_, diag_, up_, lo_, invdiag_;
};

// ---- Helmholtz (helmholtz.h/.cpp) ------------------------------------------
class HelmholtzSolver {
   public:
    HelmholtzSolver() = default;
    HelmholtzSolver(int numberModes, Real a, Real b, Real lambda, Real nu = 1.0);  // helmholtz.cpp:18
    void solve(ChebyCoeff& u, const ChebyCoeff& f, Real ua, Real ub) const;       // :79
    void solve(ChebyCoeff& u, Real& mu, const ChebyCoeff& f, Real umean, Real ua, Real ub) const;  // :158
    const BandedTridiag& Ae() const { return Ae_; }
    const BandedTridiag& Ao() const { return Ao_; }

   private:
    int beta(int n) const { return (n > N_ - 2) ? 0 : 1; }    // helmholtz.h:65
    int c(int n) const { return (n == 0 || n == N_) ? 2 : 1; }  // helmholtz.h:66
    int N_ = 0, nModes_ = 0, nEven_ = 0, nOdd_ = 0;
    Real a_ = 0, b_ = 0, lambda_ = 0, nu_ = 0;
    BandedTridiag Ae_, Ao_, Be_, Bo_;
};

// ---- TauSolver (tausolver.h/.cpp) ------------------------------------------
class TauSolver {
   public:
    TauSolver() = default;
    TauSolver(int kx, int kz, Real Lx, Real Lz, Real a, Real b, Real lambda, Real nu, int Ny, bool tauCorrection);
    void solve(ComplexChebyCoeff& u, ComplexChebyCoeff& v, ComplexChebyCoeff& w, ComplexChebyCoeff& P,
               const ComplexChebyCoeff& Rx, const ComplexChebyCoeff& Ry, const ComplexChebyCoeff& Rz) const;
    void solve(ComplexChebyCoeff& u, ComplexChebyCoeff& v, ComplexChebyCoeff& w, ComplexChebyCoeff& P, Real& dPdx,
               Real& dPdz, const ComplexChebyCoeff& Rx, const ComplexChebyCoeff& Ry, const ComplexChebyCoeff& Rz,
               Real umean, Real wmean) const;
    void influenceCorrection(ChebyCoeff& P, ChebyCoeff& v) const;
    void solve_P_and_v(ChebyCoeff& P, ChebyCoeff& v, const ChebyCoeff& r, const ChebyCoeff& Ry, Real& sigmaNb1,
                       Real& sigmaNb) const;

   private:
    int N_ = 0, Nb_ = 0, kx_ = 0, kz_ = 0;
    Real two_pi_kxLx_ = 0, two_pi_kzLz_ = 0, kappa2_ = 0, a_ = 0, b_ = 0, lambda_ = 0, nu_ = 0;
    bool tauCorrection_ = true;
    HelmholtzSolver pressureHelmholtz_, velocityHelmholtz_;
    ChebyCoeff P_0_, v_0_, P_plus_, v_plus_, P_minus_, v_minus_;
    Real i00_ = 0, i01_ = 0, i10_ = 0, i11_ = 0, sigma0_Nb1_ = 0, sigma0_Nb_ = 0;
};

// ---- FlowField (flowfield.h/.cpp, serial layout) ----------------------------
class FlowField {
   public:
    FlowField() = default;
    FlowField(int Nx, int Ny, int Nz, int Nd, Real Lx, Real Lz, Real a, Real b, FieldState xz = Spectral,
              FieldState y = Spectral);
    void resize(int Nx, int Ny, int Nz, int Nd, Real Lx, Real Lz, Real a, Real b);
    bool geomCongruent(const FlowField& f) const;
    // flowfield.h:372-402 (serial branch)
    Real& rv(int nx, int ny, int nz, int i) { return rdata_[nz + Nzpad_ * (nx + Nx_ * (ny + Ny_ * i))]; }
    Real rv(int nx, int ny, int nz, int i) const { return rdata_[nz + Nzpad_ * (nx + Nx_ * (ny + Ny_ * i))]; }
    Complex* cdata() { return reinterpret_cast<Complex*>(rdata_.data()); }
    const Complex* cdata() const { return reinterpret_cast<const Complex*>(rdata_.data()); }
    Complex& cmplx(int mx, int my, int mz, int i) { return cdata()[mz + Mz_ * (mx + Nx_ * (my + Ny_ * i))]; }
    const Complex& cmplx(int mx, int my, int mz, int i) const {
        return cdata()[mz + Mz_ * (mx + Nx_ * (my + Ny_ * i))];
    }
    Real* rdata() { return rdata_.data(); }
    const Real* rdata() const { return rdata_.data(); }
    long Nloc() const { return (long)rdata_.size(); }

    void makeSpectral_xz();  // flowfield.cpp:1850
    void makePhysical_xz();  // :1870
    void makeSpectral_y();   // :1888
    void makePhysical_y();   // :1939
    void makeSpectral() { makeSpectral_xz(); makeSpectral_y(); }  // :1989
    void makePhysical() { makePhysical_y(); makePhysical_xz(); }  // :1993
    void makeState(FieldState xz, FieldState y);
    void setState(FieldState xz, FieldState y) { xzstate_ = xz; ystate_ = y; }
    void setToZero() { std::fill(rdata_.begin(), rdata_.end(), 0.0); }
    void zeroPaddedModes();  // :2235
    void add(Real a, const FlowField& u);                            // flowfield.h:596
    void add(Real a, const FlowField& u, Real b, const FlowField& v);  // flowfield.h:606
    void scale(Real c) { for (auto& x : rdata_) x *= c; }

    int Nx() const { return Nx_; }
    int Ny() const { return Ny_; }
    int Nz() const { return Nz_; }
    int Nd() const { return Nd_; }
    int Mx() const { return Nx_; }
    int Mz() const { return Mz_; }
    int Nzpad() const { return Nzpad_; }
    Real Lx() const { return Lx_; }
    Real Lz() const { return Lz_; }
    Real a() const { return a_; }
    Real b() const { return b_; }
    int kx(int mx) const { return (mx <= Nx_ / 2) ? mx : mx - Nx_; }  // flowfield.h:462
    int mx(int kx) const { return (kx >= 0) ? kx : kx + Nx_; }
    int kxmax() const { return Nx_ / 2; }
    int kzmax() const { return Nz_ / 2; }
    int kxmin() const { return Nx_ / 2 + 1 - Nx_; }
    int kxmaxDealiased() const { return Nx_ / 3 - 1; }
    int kzmaxDealiased() const { return Nz_ / 3 - 1; }
    bool isAliased(int kx, int kz) const {
        return (std::abs(kx) > kxmaxDealiased() || std::abs(kz) > kzmaxDealiased());
    }
    bool padded() const { return padded_; }
    void setPadded(bool p) { padded_ = p; }
    FieldState xzstate() const { return xzstate_; }
    FieldState ystate() const { return ystate_; }

   private:
    int Nx_ = 0, Ny_ = 0, Nz_ = 0, Nd_ = 0, Nzpad_ = 0, Mz_ = 0;
    Real Lx_ = 0, Lz_ = 0, a_ = 0, b_ = 0;
    bool padded_ = false;
    FieldState xzstate_ = Spectral, ystate_ = Spectral;
    std::vector<Real> rdata_;
};

inline int zero_last_mode(int k, int kmax, int n) { return ((k == kmax) && (n % 2 == 1)) ? 0 : 1; }  // flowfield.h:593
inline int i3j(int i, int j) { return 3 * i + j; }                                                       // basisfunc.h:217

Real L2Norm2(const FlowField& u, bool normalize = true);                       // diffops.cpp:425
Real L2Dist2(const FlowField& u, const FlowField& v, bool normalize = true);    // diffops.cpp:353

// nonlinear terms (diffops.cpp)
void curl(FlowField& f, FlowField& curlf);                           // :2229
void grad(FlowField& f, FlowField& gradf);                           // :1784
void cross(FlowField& f, FlowField& g, FlowField& fcg, FieldState finalstate);  // :2560
void outer(FlowField& f, FlowField& g, FlowField& fg);               // :2336
void div(FlowField& f, FlowField& divf, FieldState finalstate);      // :2470
void rotationalNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState finalstate);     // :2852
void convectionNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState finalstate);     // :2883
void divergenceNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState finalstate);     // :3110
void skewsymmetricNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState finalstate);  // :3142
void dotgrad(FlowField& u, FlowField& v, FlowField& f, FlowField& tmp);                   // :3586
void navierstokesNL(FlowField& u, const ChebyCoeff& Ubase, const ChebyCoeff& Wbase, FlowField& f, FlowField& tmp,
                    DNSFlags& flags);  // nse.cpp:12

ChebyCoeff laminarProfile(Real nu, MeanConstraint constraint, Real dPdx, Real Ubulk, Real Vsuck, Real a, Real b,
                          Real ua, Real ub, int Ny);  // nse.cpp:823

// PressureSolver(u, U, W, nu, Vsuck, nonl).solve(p, u) (poissonsolver.cpp:270-431)
void pressure_solve(const FlowField& u, const ChebyCoeff& Ubase, const ChebyCoeff& Wbase, Real nu, Real Vsuck,
                    NonlinearMethod nonl, FlowField& p);

// ---- NSE (nse.h/.cpp) -------------------------------------------------------
class NSE {
   public:
    NSE(const std::vector<FlowField>& fields, const DNSFlags& flags);                                     // nse.cpp:212
    NSE(const std::vector<FlowField>& fields, const std::vector<ChebyCoeff>& base, const DNSFlags& flags);  // nse.cpp:291
    void nonlinear(std::vector<FlowField>& in, std::vector<FlowField>& out);                // nse.cpp:383
    void linear(const std::vector<FlowField>& in, std::vector<FlowField>& out);             // nse.cpp:393
    void solve(std::vector<FlowField>& out, const std::vector<FlowField>& rhs, int s = 0);  // nse.cpp:479
    void reset_lambda(const std::vector<Real>& lambda_t);                                  // nse.cpp:673
    void reset_gradp(Real dPdx, Real dPdz);
    void reset_bulkv(Real Ubulk, Real Wbulk);
    bool isAliasedMode(int kx, int kz) const { return std::abs(kx) > kxd_max_ || std::abs(kz) > kzd_max_; }
    const ChebyCoeff& Ubase() const { return Ubase_; }
    const ChebyCoeff& Wbase() const { return Wbase_; }
    Real dPdx() const { return dPdxAct_; }
    Real Ubulk() const { return UbulkAct_; }
    Real dPdxRef() const { return dPdxRef_; }
    Real UbulkRef() const { return UbulkRef_; }
    DNSFlags& flags() { return flags_; }

   private:
    void setup(const std::vector<FlowField>& fields);
    void createCFBaseFlow();
    void initCFConstraint(const FlowField& u);
    std::vector<Real> lambda_t_;
    std::vector<std::vector<TauSolver>> tausolver_;  // [substep][mx*Mz+mz]
    DNSFlags flags_;
    int Nx_ = 0, Nz_ = 0, Mx_ = 0, My_ = 0, Mz_ = 0, Nyd_ = 0, kxd_max_ = 0, kzd_max_ = 0, kxmax_ = 0, kzmax_ = 0;
    Real Lx_ = 0, Lz_ = 0, a_ = 0, b_ = 0;
    Real dPdxRef_ = 0, dPdxAct_ = 0, dPdzRef_ = 0, dPdzAct_ = 0, UbulkRef_ = 0, UbulkAct_ = 0, UbulkBase_ = 0;
    Real WbulkRef_ = 0, WbulkAct_ = 0, WbulkBase_ = 0;
    ChebyCoeff Ubase_, Ubaseyy_, Wbase_, Wbaseyy_;
    FlowField tmp_;
    ComplexChebyCoeff uk_, vk_, wk_, Pk_, Pyk_, Ruk_, Rvk_, Rwk_;
};

// ---- time-stepping algorithms (dnsalgo.h/.cpp) -------------------------------
class DNSAlgorithm {
   public:
    DNSAlgorithm(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags);
    virtual ~DNSAlgorithm() = default;
    virtual void advance(std::vector<FlowField>& fields, int Nsteps) = 0;
    virtual bool push(std::vector<FlowField>& fields) { return true; }
    virtual bool full() const { return true; }
    virtual void reset_dt(Real dt) = 0;
    Real time() const { return t_; }
    void reset_time(Real t) { t_ = t; }
    int Ninitsteps() const { return Ninitsteps_; }
    Real CFL(FlowField& u) const;  // dnsalgo.cpp:63

   protected:
    DNSFlags flags_;
    int order_ = 0, numfields_ = 0, Ninitsteps_ = 0;
    Real t_ = 0;
    std::vector<Real> lambda_t_;
    std::shared_ptr<NSE> nse_;
};

class MultistepDNS : public DNSAlgorithm {  // dnsalgo.cpp:102
   public:
    MultistepDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags);
    void advance(std::vector<FlowField>& fields, int Nsteps) override;  // :202
    bool push(std::vector<FlowField>& fields) override;          // :287
    bool full() const override { return countdown_ == 0; }
    void reset_dt(Real dt) override;

   private:
    Real eta_ = 0;
    std::vector<Real> alpha_, beta_;
    std::vector<std::vector<FlowField>> fields_, nonlf_;
    int countdown_ = 0;
};

class CNABstyleDNS : public DNSAlgorithm {  // dnsalgo.cpp:510
   public:
    CNABstyleDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags);
    void advance(std::vector<FlowField>& fields, int Nsteps) override;  // :641
    bool push(std::vector<FlowField>& fields) override;          // :598
    bool full() const override { return full_; }
    void reset_dt(Real dt) override;

   private:
    int Nsubsteps_ = 0;
    bool full_ = false;
    std::vector<FlowField> fj1_, fj_;
    std::vector<Real> alpha_, beta_, gamma_, zeta_;
};

class RungeKuttaDNS : public DNSAlgorithm {  // dnsalgo.cpp:351
   public:
    RungeKuttaDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags);
    void advance(std::vector<FlowField>& fields, int Nsteps) override;  // :416
    void reset_dt(Real dt) override;

   private:
    int Nsubsteps_ = 0;
    std::vector<FlowField> Qj1_, Qj_;
    std::vector<Real> A_, B_, C_;
};

class DNS {  // dns.h / dns.cpp
   public:
    DNS(const std::vector<FlowField>& fields, const DNSFlags& flags);  // dns.cpp:22
    DNS(const std::vector<FlowField>& fields, const std::vector<ChebyCoeff>& base, const DNSFlags& flags);  // :41
    void advance(std::vector<FlowField>& fields, int nSteps = 1);  // dns.cpp:135
    void reset_dt(Real dt);
    Real time() const { return main_->time(); }
    Real dPdx() const { return main_nse_->dPdx(); }
    Real Ubulk() const { return main_nse_->Ubulk(); }
    Real CFL(FlowField& u) const { return main_->CFL(u); }
    const ChebyCoeff& Ubase() const { return main_nse_->Ubase(); }
    const ChebyCoeff& Wbase() const { return main_nse_->Wbase(); }
    void reset_gradp(Real dPdx, Real dPdz);
    void reset_bulkv(Real Ubulk, Real Wbulk);

   private:
    static std::unique_ptr<DNSAlgorithm> newAlgorithm(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse,
                                                      const DNSFlags& flags);
    std::shared_ptr<NSE> main_nse_, init_nse_;
    std::unique_ptr<DNSAlgorithm> main_, init_;
};

}  // namespace cfref
