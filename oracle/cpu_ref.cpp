// cpu_ref -- FP64 CPU restatement of channelflow's DNS time step.
// TEST INFRASTRUCTURE ONLY (the oracle): see cpu_ref.hpp for the rules.
// Citations are channelflow 2.0.2 source paths, file:line.
#include "cpu_ref.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

namespace cfref {

static inline Real square(Real x) { return x * x; }

// ============================================================================
// FFTs. A recursive mixed-radix decimation-in-time DFT with twiddles evaluated
// in long double; radix 4 and 2 specialised, any other prime by a direct DFT.
// ============================================================================
static std::vector<int> factorize(int n) {
    std::vector<int> f;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (f.empty()) f.push_back(1);
    return f;
}

CFFT::CFFT(int n) : n_(n), fac_(factorize(n)), w_(n), scratch_(n) {
    for (int k = 0; k < n; ++k) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        w_[k] = Complex((double)cosl(ang), (double)sinl(ang));
    }
}

void CFFT::rec(Complex* out, const Complex* in, int nsub, int stride, int fi, int sign) const {
    const int p = fac_[fi];
    const int m = nsub / p;
    const int tstep = n_ / nsub;  // twiddle stride for this stage
    if (m == 1) {
        for (int q = 0; q < p; ++q) out[q] = in[q * stride];
    } else {
        for (int q = 0; q < p; ++q) rec(out + q * m, in + q * stride, m, stride * p, fi + 1, sign);
    }
    auto tw = [&](long idx) -> Complex {  // idx < n_ for every call below
        Complex c = w_[idx];
        return sign < 0 ? c : std::conj(c);
    };
    if (p == 1) return;
    if (p == 2) {
        for (int k = 0; k < m; ++k) {
            Complex a = out[k], b = out[k + m] * tw((long)k * tstep);
            out[k] = a + b;
            out[k + m] = a - b;
        }
        return;
    }
    if (p == 4) {
        for (int k = 0; k < m; ++k) {
            Complex x0 = out[k];
            Complex x1 = out[k + m] * tw((long)k * tstep);
            Complex x2 = out[k + 2 * m] * tw(2L * k * tstep);
            Complex x3 = out[k + 3 * m] * tw(3L * k * tstep);
            Complex s02 = x0 + x2, d02 = x0 - x2, s13 = x1 + x3, d13 = x1 - x3;
            // multiply d13 by -i (forward) or +i (backward)
            Complex jd13 = (sign < 0) ? Complex(d13.imag(), -d13.real()) : Complex(-d13.imag(), d13.real());
            out[k] = s02 + s13;
            out[k + m] = d02 + jd13;
            out[k + 2 * m] = s02 - s13;
            out[k + 3 * m] = d02 - jd13;
        }
        return;
    }
    if (p == 3) {
        const Complex w1 = tw((long)(n_ / 3));  // n_/3 < n_
        const Real s3 = w1.imag();  // -sqrt(3)/2 forward, +sqrt(3)/2 backward
        for (int k = 0; k < m; ++k) {
            Complex x0 = out[k];
            Complex x1 = out[k + m] * tw((long)k * tstep);
            Complex x2 = out[k + 2 * m] * tw(2L * k * tstep);
            Complex s12 = x1 + x2, d12 = x1 - x2;
            Complex a = x0 - 0.5 * s12;
            Complex b = Complex(-s3 * d12.imag(), s3 * d12.real());  // i*s3*d12
            out[k] = x0 + s12;
            out[k + m] = a + b;
            out[k + 2 * m] = a - b;
        }
        return;
    }
    Complex tbuf[64], ybuf[64], rbuf[64];
    std::vector<Complex> theap, yheap, rheap;
    Complex *t = tbuf, *y = ybuf, *root = rbuf;
    if (p > 64) {
        theap.resize(p); yheap.resize(p); rheap.resize(p);
        t = theap.data(); y = yheap.data(); root = rheap.data();
    }
    for (int j = 0; j < p; ++j) root[j] = tw((long)j * (n_ / p));
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) t[q] = out[k + q * m] * tw((long)q * k * tstep);
        for (int r = 0; r < p; ++r) {
            Complex s = 0;
            int idx = 0;
            for (int q = 0; q < p; ++q) {
                s += t[q] * root[idx];
                idx += r;
                if (idx >= p) idx -= p;
            }
            y[r] = s;
        }
        for (int r = 0; r < p; ++r) out[k + r * m] = y[r];
    }
}

void CFFT::exec(Complex* x, int sign) const {
    std::copy(x, x + n_, scratch_.begin());
    rec(x, scratch_.data(), n_, 1, 0, sign);
}

RFFT::RFFT(int n) : n_(n), half_(n / 2), tw_(n / 2 + 1), z_(n / 2) {
    assert(n % 2 == 0);
    for (int k = 0; k <= n / 2; ++k) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        tw_[k] = Complex((double)cosl(ang), (double)sinl(ang));
    }
}

void RFFT::r2c(const Real* in, Complex* out) const {
    const int h = n_ / 2;
    for (int j = 0; j < h; ++j) z_[j] = Complex(in[2 * j], in[2 * j + 1]);
    half_.exec(z_.data(), -1);
    for (int k = 0; k <= h; ++k) {
        Complex zk = z_[k % h];
        Complex zc = std::conj(z_[(h - k) % h]);
        Complex E = 0.5 * (zk + zc);
        Complex O = Complex(0, -0.5) * (zk - zc);
        out[k] = E + tw_[k] * O;
    }
}

void RFFT::c2r(const Complex* in, Real* out) const {
    const int h = n_ / 2;
    for (int k = 0; k < h; ++k) {
        Complex xk = in[k];
        Complex xc = std::conj(in[h - k]);
        if (k == 0) { xk = Complex(in[0].real(), 0.0); xc = Complex(in[h].real(), 0.0); }
        Complex s = xk + xc;
        Complex d = (xk - xc) * std::conj(tw_[k]);
        z_[k] = s + Complex(-d.imag(), d.real());  // s + i d
    }
    half_.exec(z_.data(), +1);
    for (int j = 0; j < h; ++j) {
        out[2 * j] = z_[j].real();
        out[2 * j + 1] = z_[j].imag();
    }
}

DCT1::DCT1(int N) : N_(N), r_(2 * (N - 1)), e_(2 * (N - 1)), E_(N) {}

void DCT1::exec(Real* x) const {
    const int M = 2 * (N_ - 1);
    for (int j = 0; j < N_; ++j) e_[j] = x[j];
    for (int j = N_; j < M; ++j) e_[j] = x[M - j];
    r_.r2c(e_.data(), E_.data());
    for (int k = 0; k < N_; ++k) x[k] = E_[k].real();
}

template <class T>
static const T& cached(int n) {
    static std::map<int, std::unique_ptr<T>> cache;
    static std::mutex mtx;
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cache.find(n);
    if (it == cache.end()) it = cache.emplace(n, std::make_unique<T>(n)).first;
    return *it->second;
}

// ============================================================================
// ChebyCoeff (chebyshev.cpp)
// ============================================================================
Real ChebyCoeff::eval_b() const {  // chebyshev.cpp:405
    if (state == Spectral) {
        Real sum = 0.0;
        for (int n = N() - 1; n >= 0; --n) sum += d[n];
        return sum;
    }
    return d[0];
}

Real ChebyCoeff::eval_a() const {  // chebyshev.cpp:418
    if (state == Spectral) {
        Real sum = 0.0;
        for (int n = N() - 1; n >= 0; --n) sum += d[n] * ((n % 2 == 0) ? 1 : -1);
        return sum;
    }
    return d[N() - 1];
}

Real ChebyCoeff::mean() const {  // chebyshev.cpp:505
    Real sum = d[0];
    for (int n = 2; n < N(); n += 2) sum -= d[n] / (double)(n * n - 1);
    return sum;
}

void ChebyCoeff::makeSpectral() {  // chebyshev.cpp:262 (chebyfft)
    if (state == Spectral) return;
    const int N_ = N();
    if (N_ >= 2) {
        cached<DCT1>(N_).exec(d.data());
        Real nrm = 1.0 / (N_ - 1);
        d[0] *= 0.5 * nrm;
        for (int n = 1; n < N_ - 1; ++n) d[n] *= nrm;
        d[N_ - 1] *= 0.5 * nrm;
    }
    state = Spectral;
}

void ChebyCoeff::makePhysical() {  // chebyshev.cpp:286 (ichebyfft)
    if (state == Physical) return;
    const int N_ = N();
    if (N_ >= 2) {
        d[0] *= 2.0;
        d[N_ - 1] *= 2.0;
        cached<DCT1>(N_).exec(d.data());
        for (auto& x : d) x *= 0.5;
    }
    state = Physical;
}

void diff(const ChebyCoeff& u, ChebyCoeff& du) {  // chebyshev.cpp:672
    const int N = u.N();
    if ((int)du.d.size() != N) du.d.assign(N, 0.0);
    du.a = u.a;
    du.b = u.b;
    du.state = Spectral;
    const int Nb = N - 1;
    if (Nb == -1) return;
    if (Nb == 0) { du[0] = 0.0; return; }
    const Real scale = 4.0 / u.L();
    du[Nb] = 0.0;
    du[Nb - 1] = scale * Nb * u[Nb];
    for (int n = Nb - 2; n >= 0; --n) du[n] = du[n + 2] + scale * (n + 1) * u[n + 1];
    du[0] *= 0.5;
}

void diff2(const ChebyCoeff& u, ChebyCoeff& d2u) {  // chebyshev.cpp:699
    ChebyCoeff tmp(u.N(), u.a, u.b);
    diff(u, tmp);
    diff(tmp, d2u);
}

void diff(const ComplexChebyCoeff& u, ComplexChebyCoeff& du) {  // chebyshev.cpp:1206
    if (du.N() != u.N()) du = ComplexChebyCoeff(u.N(), u.re.a, u.re.b);
    diff(u.re, du.re);
    diff(u.im, du.im);
}

Real L2Norm2(const ChebyCoeff& u, bool normalize) {  // chebyshev.cpp:758
    const int N = u.N();
    const double e = 1.0;
    Real sum = 0.0;
    for (int m = N - 1; m >= 0; --m) {
        Real psum = 0.0;
        for (int n = m % 2; n < N; n += 2)
            psum += (u[n] * (e - (double)m * m - (double)n * n)) /
                    ((e + m - n) * (e - m + n) * (e + m + n) * (e - m - n));
        sum += u[m] * psum;
    }
    if (!normalize) sum *= u.b - u.a;
    return sum;
}

std::vector<Real> chebypoints(int N, Real a, Real b) {  // chebyshev.cpp:746
    std::vector<Real> x(N);
    Real piN = pi / (N - 1);
    Real radius = (b - a) / 2, center = (b + a) / 2;
    for (int j = 0; j < N; ++j) x[j] = center + radius * cos(j * piN);
    return x;
}

// ============================================================================
// BandedTridiag (bandedtridiag.cpp). Row 0 is full ("band"); rows 1..M-1 are
// tridiagonal. UL decomposition without pivoting.
// ============================================================================
BandedTridiag::BandedTridiag(int M)
    : M_(M), band_(M, 0.0), diag_(M, 0.0), up_(M, 0.0), lo_(M, 0.0), invdiag_(M, 0.0) {}

void BandedTridiag::ULdecomp() {  // bandedtridiag.cpp:212-229
    const int Mb = M_ - 1;
    for (int k = Mb; k > 1; --k) {
        Real Akk = diag_[k];
        Real w = lo_[k];
        up_[k - 1] /= Akk;
        diag_[k - 1] -= w * up_[k - 1];
        band_[k] /= Akk;
        band_[k - 1] -= w * band_[k];
    }
    band_[1] /= diag_[1];
    band_[0] -= lo_[1] * band_[1];
    diag_[0] = band_[0];  // diag(0) and band(0) are one storage cell (bandedtridiag.h:79-90)
    for (int i = 0; i < M_; ++i) invdiag_[i] = 1.0 / diag_[i];
}

void BandedTridiag::ULsolveStrided(Real* b, int offset, int stride) const {  // bandedtridiag.cpp:232-277
    const int Mb = M_ - 1;
    int i;
    for (i = Mb - 1; i > 0; --i) b[offset + stride * i] -= up_[i] * b[offset + stride * (i + 1)];
    for (int j = i + 1; j < M_; ++j) b[offset] -= band_[j] * b[offset + stride * j];
    b[offset] /= diag_[0];
    for (i = 1; i < M_; ++i)
        b[offset + stride * i] = (b[offset + stride * i] - lo_[i] * b[offset + stride * (i - 1)]) * invdiag_[i];
}

void BandedTridiag::multiplyStrided(const Real* x, Real* b, int offset, int stride) const {  // :315-333
    const int Mbar = M_ - 1;
    Real sum = 0.0;
    for (int j = 0; j < M_; ++j) sum += band_[j] * x[offset + stride * j];
    b[offset] = sum;
    for (int i = 1; i < Mbar; ++i)
        b[offset + stride * i] = lo_[i] * x[offset + stride * (i - 1)] + diag_[i] * x[offset + stride * i] +
                                 up_[i] * x[offset + stride * (i + 1)];
    b[offset + stride * Mbar] = lo_[Mbar] * x[offset + stride * (Mbar - 1)] + diag_[Mbar] * x[offset + stride * Mbar];
}

// ============================================================================
// HelmholtzSolver (helmholtz.cpp): nu u'' - lambda u = f, u(a)=ua, u(b)=ub, by
// the Chebyshev-tau quasi-tridiagonal form of Canuto & Hussaini 5.1.24.
// ============================================================================
HelmholtzSolver::HelmholtzSolver(int numberModes, Real a, Real b, Real lambda, Real nu)
    : N_(numberModes - 1), nModes_(numberModes), nEven_(N_ / 2 + 1), nOdd_(N_ / 2), a_(a), b_(b),
      lambda_(lambda), nu_(nu), Ae_(nEven_), Ao_(nOdd_), Be_(nEven_), Bo_(nOdd_) {  // helmholtz.cpp:18-77
    if (nModes_ % 2 != 1 || nModes_ <= 2) throw std::runtime_error("HelmholtzSolver: Ny must be odd and > 2");
    Real nuscaled = nu / square((b_ - a_) / 2);
    Be_.diag(0) = 1.0;
    Be_.band(0) = 1.0;  // same storage cell as diag(0)
    int i;
    for (i = 0; i < nEven_; ++i) Ae_.band(i) = 1.0;
    Ae_.diag(0) = Ae_.band(0);
    for (i = 1; i < nEven_; ++i) {
        int n = 2 * i;
        int nn = n * n;
        Ae_.lodiag(i) = -(c(n - 2) * lambda) / (4 * n * (n - 1));
        Ae_.diag(i) = (nuscaled + (beta(n) * lambda) / (2 * (nn - 1)));
        if (beta(n + 2)) Ae_.updiag(i) = -lambda / (4 * n * (n + 1));
        Be_.lodiag(i) = ((Real)c(n - 2)) / (4 * n * (n - 1));
        Be_.diag(i) = -((Real)beta(n)) / (2 * (nn - 1));
        if (beta(n + 2)) Be_.updiag(i) = 1.0 / (4 * n * (n + 1));
    }
    Bo_.diag(0) = 1.0;
    Bo_.band(0) = 1.0;
    for (i = 0; i < nOdd_; ++i) Ao_.band(i) = 1.0;
    Ao_.diag(0) = Ao_.band(0);
    for (i = 1; i < nOdd_; ++i) {
        int n = 2 * i + 1;
        int nn = n * n;
        Ao_.lodiag(i) = -(c(n - 2) * lambda) / (4 * n * (n - 1));
        Ao_.diag(i) = (nuscaled + (beta(n) * lambda) / (2 * (nn - 1)));
        if (beta(n + 2)) Ao_.updiag(i) = -lambda / (4 * n * (n + 1));
        Bo_.lodiag(i) = ((Real)c(n - 2)) / (4 * n * (n - 1));
        Bo_.diag(i) = -((Real)beta(n)) / (2 * (nn - 1));
        if (beta(n + 2)) Bo_.updiag(i) = 1.0 / (4 * n * (n + 1));
    }
    Ae_.ULdecomp();
    Ao_.ULdecomp();
}

void HelmholtzSolver::solve(ChebyCoeff& u, const ChebyCoeff& f, Real ua, Real ub) const {  // helmholtz.cpp:79-95
    if (u.N() != nModes_) u = ChebyCoeff(nModes_, a_, b_);
    Be_.multiplyStrided(f.d.data(), u.d.data(), 0, 2);
    Bo_.multiplyStrided(f.d.data(), u.d.data(), 1, 2);
    u[0] = (ub + ua) / 2;
    u[1] = (ub - ua) / 2;
    Ae_.ULsolveStrided(u.d.data(), 0, 2);
    Ao_.ULsolveStrided(u.d.data(), 1, 2);
    u.state = Spectral;
}

void HelmholtzSolver::solve(ChebyCoeff& u, Real& mu, const ChebyCoeff& f, Real umean, Real ua,
                            Real ub) const {  // helmholtz.cpp:158-213
    int N = nModes_;
    ChebyCoeff utmp(N, a_, b_);
    ChebyCoeff rtmp(f);
    solve(utmp, rtmp, ua, ub);
    Real uamean = utmp.mean();
    rtmp.setToZero();
    rtmp[0] = nu_;
    solve(utmp, rtmp, 0.0, 0.0);
    Real ucmean = utmp.mean();
    mu = nu_ * (umean - uamean) / ucmean;
    rtmp = f;
    rtmp[0] += mu;
    solve(u, rtmp, ua, ub);
}

// ============================================================================
// TauSolver (tausolver.cpp): Canuto-Hussaini influence-matrix Stokes solve
// per (kx,kz) with tau correction.
// ============================================================================
static int n_func(int k, int Nb) {  // tausolver.cpp:15
    if (k == 0) return Nb - 1;
    if (k == Nb) return 0;
    if (k % 2 == 0) return 2 * (Nb - 1);
    return 2 * Nb;
}

TauSolver::TauSolver(int kx, int kz, Real Lx, Real Lz, Real a, Real b, Real lambda, Real nu, int Ny,
                     bool tauCorrection)
    : N_(Ny), Nb_(Ny - 1), kx_(kx), kz_(kz), two_pi_kxLx_(2 * pi * kx / Lx), two_pi_kzLz_(2 * pi * kz / Lz),
      kappa2_(4 * square(pi) * (square(kx / Lx) + square(kz / Lz))), a_(a), b_(b), lambda_(lambda), nu_(nu),
      tauCorrection_(tauCorrection), pressureHelmholtz_(N_, a_, b_, kappa2_),
      velocityHelmholtz_(N_, a_, b_, lambda_, nu_), P_0_(N_, a_, b_), v_0_(N_, a_, b_), P_plus_(N_, a_, b_),
      v_plus_(N_, a_, b_), P_minus_(N_, a_, b_), v_minus_(N_, a_, b_) {  // tausolver.cpp:81-176
    ChebyCoeff dvplus_dy(N_, a_, b_), dvminus_dy(N_, a_, b_), zero(N_, a_, b_), dPdy(N_, a_, b_);
    pressureHelmholtz_.solve(P_plus_, zero, 0.0, 1.0);
    diff(P_plus_, dPdy);
    velocityHelmholtz_.solve(v_plus_, dPdy, 0.0, 0.0);
    pressureHelmholtz_.solve(P_minus_, zero, 1.0, 0.0);
    diff(P_minus_, dPdy);
    velocityHelmholtz_.solve(v_minus_, dPdy, 0.0, 0.0);
    diff(v_plus_, dvplus_dy);
    diff(v_minus_, dvminus_dy);
    Real A = dvplus_dy.eval_b(), B = dvminus_dy.eval_b(), C = dvplus_dy.eval_a(), D = dvminus_dy.eval_a();
    Real discriminant = A * D - B * C;
    i00_ = D / discriminant;
    i01_ = -B / discriminant;
    i10_ = -C / discriminant;
    i11_ = A / discriminant;
    ChebyCoeff p0_rhs(N_, a_, b_);
    Real c = 2 / (b_ - a_);
    for (int i = 0; i <= Nb_; ++i) p0_rhs[i] = c * n_func(i, Nb_);
    pressureHelmholtz_.solve(P_0_, p0_rhs, 0.0, 0.0);
    ChebyCoeff dP0dy(N_, a_, b_);
    diff(P_0_, dP0dy);
    velocityHelmholtz_.solve(v_0_, dP0dy, 0.0, 0.0);
    influenceCorrection(P_0_, v_0_);
    ChebyCoeff v0yy(N_, a_, b_);
    diff2(v_0_, v0yy);
    sigma0_Nb1_ = lambda_ * v_0_[Nb_ - 1] + dP0dy[Nb_ - 1] - nu_ * v0yy[Nb_ - 1];
    sigma0_Nb_ = lambda_ * v_0_[Nb_] + dP0dy[Nb_] - nu_ * v0yy[Nb_];
}

void TauSolver::influenceCorrection(ChebyCoeff& P, ChebyCoeff& v) const {  // tausolver.cpp:178-191
    ChebyCoeff tmp(N_, a_, b_);
    diff(v, tmp);
    Real dvp_dy_plus = tmp.eval_b();
    Real dvp_dy_minus = tmp.eval_a();
    Real delta_plus = -i00_ * dvp_dy_plus - i01_ * dvp_dy_minus;
    Real delta_minus = -i10_ * dvp_dy_plus - i11_ * dvp_dy_minus;
    for (int i = 0; i < N_; ++i) {
        P[i] += delta_plus * P_plus_[i] + delta_minus * P_minus_[i];
        v[i] += delta_plus * v_plus_[i] + delta_minus * v_minus_[i];
    }
}

void TauSolver::solve_P_and_v(ChebyCoeff& P, ChebyCoeff& v, const ChebyCoeff& r, const ChebyCoeff& Ry,
                              Real& sigmaNb1, Real& sigmaNb) const {  // tausolver.cpp:193-251
    pressureHelmholtz_.solve(P, r, 0.0, 0.0);
    if (kx_ == 0 && kz_ == 0) {
        for (int i = 0; i < N_; ++i) v[i] = 0.0;
        return;
    }
    ChebyCoeff tmp(N_, a_, b_);
    diff(P, tmp);
    for (int i = 0; i < N_; ++i) tmp[i] -= Ry[i];
    velocityHelmholtz_.solve(v, tmp, 0.0, 0.0);
    influenceCorrection(P, v);
    if (!tauCorrection_) return;
    diff2(v, tmp);
    Real sigma1_Nb = lambda_ * v[Nb_] - nu_ * tmp[Nb_] - Ry[Nb_];
    Real sigma1_Nb1 = lambda_ * v[Nb_ - 1] - nu_ * tmp[Nb_ - 1] - Ry[Nb_ - 1];
    diff(P, tmp);
    sigma1_Nb += tmp[Nb_];
    sigma1_Nb1 += tmp[Nb_ - 1];
    sigmaNb = sigma1_Nb / (1.0 - sigma0_Nb_);
    sigmaNb1 = sigma1_Nb1 / (1.0 - sigma0_Nb1_);
    for (int i = 0; i <= Nb_; ++i) {
        P[i] += ((i % 2 == 0) ? sigmaNb1 : sigmaNb) * P_0_[i];
        v[i] += ((i % 2 == 0) ? sigmaNb : sigmaNb1) * v_0_[i];
    }
}

static inline Complex mulI(Real a, Complex p) {  // (a*I)*p as std::complex evaluates it
    return Complex(0.0, a) * p;
}

void TauSolver::solve(ComplexChebyCoeff& u, ComplexChebyCoeff& v, ComplexChebyCoeff& w, ComplexChebyCoeff& P,
                      const ComplexChebyCoeff& Rx, const ComplexChebyCoeff& Ry,
                      const ComplexChebyCoeff& Rz) const {  // tausolver.cpp:347-402
    ComplexChebyCoeff r(N_, a_, b_);
    Real sigmaNb1, sigmaNb;
    ChebyCoeff rr(N_, a_, b_);
    diff(Ry.re, rr);
    for (int n = 0; n < N_; ++n) rr[n] -= two_pi_kxLx_ * Rx.im[n] + two_pi_kzLz_ * Rz.im[n];
    solve_P_and_v(P.re, v.re, rr, Ry.re, sigmaNb1, sigmaNb);
    diff(Ry.im, rr);
    for (int n = 0; n < N_; ++n) rr[n] += two_pi_kxLx_ * Rx.re[n] + two_pi_kzLz_ * Rz.re[n];
    solve_P_and_v(P.im, v.im, rr, Ry.im, sigmaNb1, sigmaNb);
    for (int n = 0; n < N_; ++n) r.set(n, mulI(two_pi_kxLx_, P.get(n)) - Rx.get(n));
    velocityHelmholtz_.solve(u.re, r.re, 0.0, 0.0);
    velocityHelmholtz_.solve(u.im, r.im, 0.0, 0.0);
    for (int n = 0; n < N_; ++n) r.set(n, mulI(two_pi_kzLz_, P.get(n)) - Rz.get(n));
    velocityHelmholtz_.solve(w.re, r.re, 0.0, 0.0);
    velocityHelmholtz_.solve(w.im, r.im, 0.0, 0.0);
}

void TauSolver::solve(ComplexChebyCoeff& u, ComplexChebyCoeff& v, ComplexChebyCoeff& w, ComplexChebyCoeff& P,
                      Real& dPdx, Real& dPdz, const ComplexChebyCoeff& Rx, const ComplexChebyCoeff& Ry,
                      const ComplexChebyCoeff& Rz, Real umean, Real wmean) const {  // tausolver.cpp:404-450
    ComplexChebyCoeff r(N_, a_, b_);
    Real sigmaNb1, sigmaNb;
    ChebyCoeff rr(N_, a_, b_);
    diff(Ry.re, rr);
    for (int n = 0; n < N_; ++n) rr[n] -= two_pi_kxLx_ * Rx.im[n] + two_pi_kzLz_ * Rz.im[n];
    solve_P_and_v(P.re, v.re, rr, Ry.re, sigmaNb1, sigmaNb);
    diff(Ry.im, rr);
    for (int n = 0; n < N_; ++n) rr[n] += two_pi_kxLx_ * Rx.re[n] + two_pi_kzLz_ * Rz.re[n];
    solve_P_and_v(P.im, v.im, rr, Ry.im, sigmaNb1, sigmaNb);
    for (int n = 0; n < N_; ++n) r.set(n, mulI(two_pi_kxLx_, P.get(n)) - Rx.get(n));
    velocityHelmholtz_.solve(u.re, dPdx, r.re, umean, 0.0, 0.0);
    velocityHelmholtz_.solve(u.im, r.im, 0.0, 0.0);
    for (int n = 0; n < N_; ++n) r.set(n, mulI(two_pi_kzLz_, P.get(n)) - Rz.get(n));
    velocityHelmholtz_.solve(w.re, dPdz, r.re, wmean, 0.0, 0.0);
    velocityHelmholtz_.solve(w.im, r.im, 0.0, 0.0);
}

// ============================================================================
// FlowField (flowfield.cpp, serial layout)
// ============================================================================
FlowField::FlowField(int Nx, int Ny, int Nz, int Nd, Real Lx, Real Lz, Real a, Real b, FieldState xz,
                     FieldState y) {
    resize(Nx, Ny, Nz, Nd, Lx, Lz, a, b);
    xzstate_ = xz;
    ystate_ = y;
}

void FlowField::resize(int Nx, int Ny, int Nz, int Nd, Real Lx, Real Lz, Real a, Real b) {  // flowfield.cpp:466
    Nx_ = Nx; Ny_ = Ny; Nz_ = Nz; Nd_ = Nd;
    Lx_ = Lx; Lz_ = Lz; a_ = a; b_ = b;
    Nzpad_ = 2 * (Nz / 2 + 1);
    Mz_ = Nz / 2 + 1;
    rdata_.assign((size_t)Nx * Ny * Nzpad_ * Nd, 0.0);
}

bool FlowField::geomCongruent(const FlowField& f) const {
    return Nx_ == f.Nx_ && Ny_ == f.Ny_ && Nz_ == f.Nz_ && Lx_ == f.Lx_ && Lz_ == f.Lz_ && a_ == f.a_ && b_ == f.b_;
}

void FlowField::makeSpectral_xz() {  // flowfield.cpp:1850-1868 (FFTW r2c, rank {Nx,Nz}, scale 1/(Nx Nz))
    if (xzstate_ == Spectral) return;
    const RFFT& rz = cached<RFFT>(Nz_);
    const CFFT& cx = cached<CFFT>(Nx_);
    std::vector<Real> row(Nz_);
    std::vector<Complex> col(Nx_);
    for (int p = 0; p < Ny_ * Nd_; ++p) {
        Real* plane = rdata_.data() + (size_t)p * Nx_ * Nzpad_;
        Complex* cplane = reinterpret_cast<Complex*>(plane);
        for (int nx = 0; nx < Nx_; ++nx) {
            std::copy(plane + nx * Nzpad_, plane + nx * Nzpad_ + Nz_, row.begin());
            rz.r2c(row.data(), cplane + nx * Mz_);
        }
        for (int mz = 0; mz < Mz_; ++mz) {
            for (int mx = 0; mx < Nx_; ++mx) col[mx] = cplane[mz + Mz_ * mx];
            cx.exec(col.data(), -1);
            for (int mx = 0; mx < Nx_; ++mx) cplane[mz + Mz_ * mx] = col[mx];
        }
    }
    const Real scale = 1.0 / (Nx_ * Nz_);
    for (auto& x : rdata_) x *= scale;
    xzstate_ = Spectral;
}

void FlowField::makePhysical_xz() {  // flowfield.cpp:1870-1886 (FFTW c2r)
    if (xzstate_ == Physical) return;
    const RFFT& rz = cached<RFFT>(Nz_);
    const CFFT& cx = cached<CFFT>(Nx_);
    std::vector<Real> row(Nz_);
    std::vector<Complex> col(Nx_), crow(Mz_);
    for (int p = 0; p < Ny_ * Nd_; ++p) {
        Real* plane = rdata_.data() + (size_t)p * Nx_ * Nzpad_;
        Complex* cplane = reinterpret_cast<Complex*>(plane);
        for (int mz = 0; mz < Mz_; ++mz) {
            for (int mx = 0; mx < Nx_; ++mx) col[mx] = cplane[mz + Mz_ * mx];
            cx.exec(col.data(), +1);
            for (int mx = 0; mx < Nx_; ++mx) cplane[mz + Mz_ * mx] = col[mx];
        }
        for (int nx = 0; nx < Nx_; ++nx) {
            std::copy(cplane + nx * Mz_, cplane + (nx + 1) * Mz_, crow.begin());
            rz.c2r(crow.data(), row.data());
            std::copy(row.begin(), row.end(), plane + nx * Nzpad_);
        }
    }
    xzstate_ = Physical;
}

void FlowField::makeSpectral_y() {  // flowfield.cpp:1888-1937 (serial branch)
    if (ystate_ == Spectral) return;
    if (Ny_ < 2) { ystate_ = Spectral; return; }
    const DCT1& t = cached<DCT1>(Ny_);
    const Real nrm = 1.0 / (Ny_ - 1);
    std::vector<Real> s(Ny_);
    for (int i = 0; i < Nd_; ++i)
        for (int nx = 0; nx < Nx_; ++nx)
            for (int nz = 0; nz < Nzpad_; ++nz) {
                for (int ny = 0; ny < Ny_; ++ny) s[ny] = rv(nx, ny, nz, i);
                t.exec(s.data());
                rv(nx, 0, nz, i) = 0.5 * nrm * s[0];
                for (int ny = 1; ny < Ny_ - 1; ++ny) rv(nx, ny, nz, i) = nrm * s[ny];
                rv(nx, Ny_ - 1, nz, i) = 0.5 * nrm * s[Ny_ - 1];
            }
    ystate_ = Spectral;
}

void FlowField::makePhysical_y() {  // flowfield.cpp:1939-1987 (serial branch)
    if (ystate_ == Physical) return;
    if (Ny_ < 2) { ystate_ = Physical; return; }
    const DCT1& t = cached<DCT1>(Ny_);
    std::vector<Real> s(Ny_);
    for (int i = 0; i < Nd_; ++i)
        for (int nx = 0; nx < Nx_; ++nx)
            for (int nz = 0; nz < Nzpad_; ++nz) {
                s[0] = rv(nx, 0, nz, i);
                for (int ny = 1; ny < Ny_ - 1; ++ny) s[ny] = 0.5 * rv(nx, ny, nz, i);
                s[Ny_ - 1] = rv(nx, Ny_ - 1, nz, i);
                t.exec(s.data());
                for (int ny = 0; ny < Ny_; ++ny) rv(nx, ny, nz, i) = s[ny];
            }
    ystate_ = Physical;
}

void FlowField::makeState(FieldState xz, FieldState y) {  // flowfield.cpp:1999-2010 (MPI-safe order)
    if (xzstate_ == Physical && xz == Spectral) makeSpectral_xz();
    if (ystate_ == Physical && y == Spectral) makeSpectral_y();
    if (ystate_ == Spectral && y == Physical) makePhysical_y();
    if (xzstate_ == Spectral && xz == Physical) makePhysical_xz();
}

void FlowField::zeroPaddedModes() {  // flowfield.cpp:2235-2255
    FieldState xzs = xzstate_, ys = ystate_;
    makeSpectral_xz();
    for (int mx = 0; mx < Nx_; ++mx) {
        int kx_ = kx(mx);
        for (int mz = 0; mz < Mz_; ++mz)
            if (isAliased(kx_, mz))
                for (int i = 0; i < Nd_; ++i)
                    for (int ny = 0; ny < Ny_; ++ny) cmplx(mx, ny, mz, i) = 0.0;
    }
    padded_ = true;
    makeState(xzs, ys);
}

void FlowField::add(Real a, const FlowField& u) {  // flowfield.h:596-604
    const Real* ud = u.rdata_.data();
    Real* d = rdata_.data();
    const long n = Nloc();
    for (long i = 0; i < n; ++i) d[i] += a * ud[i];
}

void FlowField::add(Real a, const FlowField& u, Real b, const FlowField& v) {  // flowfield.h:606-615
    const Real* ud = u.rdata_.data();
    const Real* vd = v.rdata_.data();
    Real* d = rdata_.data();
    const long n = Nloc();
    for (long i = 0; i < n; ++i) d[i] += a * ud[i] + b * vd[i];
}

// volume-normalised L2 norms over kx in [kxmin,kxmax], kz >= 0 with weight 2
// for kz > 0 (diffops.cpp:353-487)
static Real l2sum(const FlowField& u, const FlowField* v, bool padded, bool normalize) {
    int kxmin = padded ? -u.kxmaxDealiased() : u.kxmin();
    int kxmax = padded ? u.kxmaxDealiased() : u.kxmax();
    int kzmax = padded ? u.kzmaxDealiased() : u.kzmax();
    Real sum = 0.0;
    ComplexChebyCoeff prof(u.Ny(), u.a(), u.b());
    for (int i = 0; i < u.Nd(); ++i)
        for (int kx = kxmin; kx <= kxmax; ++kx) {
            int mx = u.mx(kx);
            int cz = 1;
            for (int kz = 0; kz <= kzmax; ++kz) {
                for (int ny = 0; ny < u.Ny(); ++ny)
                    prof.set(ny, v ? u.cmplx(mx, ny, kz, i) - v->cmplx(mx, ny, kz, i) : u.cmplx(mx, ny, kz, i));
                sum += cz * (L2Norm2(prof.re, normalize) + L2Norm2(prof.im, normalize));
                cz = 2;
            }
        }
    if (!normalize) sum *= u.Lx() * u.Lz();
    return sum;
}

Real L2Norm2(const FlowField& u, bool normalize) { return l2sum(u, nullptr, u.padded(), normalize); }
Real L2Dist2(const FlowField& u, const FlowField& v, bool normalize) {
    return l2sum(u, &v, u.padded() && v.padded(), normalize);
}

// ============================================================================
// Spectral derivatives and nonlinear terms (diffops.cpp)
// ============================================================================
void curl(FlowField& f, FlowField& curlf) {  // diffops.cpp:2229-2334
    FieldState sxz = f.xzstate(), sy = f.ystate();
    f.makeSpectral();
    if (!f.geomCongruent(curlf) || curlf.Nd() != 3)
        curlf.resize(f.Nx(), f.Ny(), f.Nz(), 3, f.Lx(), f.Lz(), f.a(), f.b());
    curlf.setState(Spectral, Spectral);
    const int My = f.Ny(), Myb = My - 1, Mx = f.Mx(), Mz = f.Mz();
    const Real scale = 4.0 / (f.b() - f.a());
    const Complex zero(0.0, 0.0);
    for (int mx = 0; mx < Mx; ++mx)
        for (int mz = 0; mz < Mz; ++mz) {
            curlf.cmplx(mx, Myb, mz, 0) = zero;
            curlf.cmplx(mx, Myb - 1, mz, 0) = (Real)Myb * scale * f.cmplx(mx, Myb, mz, 2);
        }
    for (int my = Myb - 2; my >= 0; --my)
        for (int mx = 0; mx < Mx; ++mx)
            for (int mz = 0; mz < Mz; ++mz)
                curlf.cmplx(mx, my, mz, 0) =
                    curlf.cmplx(mx, my + 2, mz, 0) + (Real)(my + 1) * scale * f.cmplx(mx, my + 1, mz, 2);
    for (int mx = 0; mx < Mx; ++mx)
        for (int mz = 0; mz < Mz; ++mz) curlf.cmplx(mx, 0, mz, 0) *= 0.5;
    for (int mx = 0; mx < Mx; ++mx)
        for (int mz = 0; mz < Mz; ++mz) {
            curlf.cmplx(mx, Myb, mz, 2) = zero;
            curlf.cmplx(mx, Myb - 1, mz, 2) = -(Real)Myb * scale * f.cmplx(mx, Myb, mz, 0);
        }
    for (int my = Myb - 2; my >= 0; --my)
        for (int mx = 0; mx < Mx; ++mx)
            for (int mz = 0; mz < Mz; ++mz)
                curlf.cmplx(mx, my, mz, 2) =
                    curlf.cmplx(mx, my + 2, mz, 2) - (Real)(my + 1) * scale * f.cmplx(mx, my + 1, mz, 0);
    for (int mx = 0; mx < Mx; ++mx)
        for (int mz = 0; mz < Mz; ++mz) curlf.cmplx(mx, 0, mz, 2) *= 0.5;
    const int kxmax = f.kxmax(), kzmax = f.kzmax();
    for (int my = 0; my < My; ++my)
        for (int mx = 0; mx < Mx; ++mx) {
            int kx = f.kx(mx);
            Complex d_dx(0.0, 2.0 * pi * kx / f.Lx() * zero_last_mode(kx, kxmax, 1));
            for (int mz = 0; mz < Mz; ++mz) {
                int kz = mz;
                Complex d_dz(0.0, 2.0 * pi * kz / f.Lz() * zero_last_mode(kz, kzmax, 1));
                Complex f0 = f.cmplx(mx, my, mz, 0), f1 = f.cmplx(mx, my, mz, 1), f2 = f.cmplx(mx, my, mz, 2);
                curlf.cmplx(mx, my, mz, 0) -= d_dz * f1;
                curlf.cmplx(mx, my, mz, 1) = d_dz * f0 - d_dx * f2;
                curlf.cmplx(mx, my, mz, 2) += d_dx * f1;
            }
        }
    f.makeState(sxz, sy);
}

static void ydiff_profile(const FlowField& f, int mx, int mz, int i, ComplexChebyCoeff& g, ComplexChebyCoeff& gy) {
    for (int my = 0; my < f.Ny(); ++my) g.set(my, f.cmplx(mx, my, mz, i));
    diff(g, gy);
}

void grad(FlowField& f, FlowField& gradf) {  // diffops.cpp:1784-1941 (vector branch)
    FieldState sxz = f.xzstate(), sy = f.ystate();
    f.makeSpectral();
    if (f.Nd() != 3) throw std::runtime_error("grad: only 3d fields are on the DNS path");
    if (!f.geomCongruent(gradf) || gradf.Nd() != 9)
        gradf.resize(f.Nx(), f.Ny(), f.Nz(), 9, f.Lx(), f.Lz(), f.a(), f.b());
    else
        gradf.setToZero();
    gradf.setState(Spectral, Spectral);
    const int My = f.Ny(), Mx = f.Mx(), Mz = f.Mz(), kxmax = f.kxmax(), kzmax = f.kzmax();
    for (int i = 0; i < 3; ++i)
        for (int my = 0; my < My; ++my)
            for (int mx = 0; mx < Mx; ++mx) {
                int kx = f.kx(mx);
                Complex cxkx(0.0, 2.0 * pi * kx / f.Lx() * zero_last_mode(kx, kxmax, 1));
                for (int mz = 0; mz < Mz; ++mz) {
                    Complex czkz(0.0, 2.0 * pi * mz / f.Lz() * zero_last_mode(mz, kzmax, 1));
                    Complex fval = f.cmplx(mx, my, mz, i);
                    gradf.cmplx(mx, my, mz, i3j(i, 0)) = fval * cxkx;
                    gradf.cmplx(mx, my, mz, i3j(i, 2)) = fval * czkz;
                }
            }
    ComplexChebyCoeff g(My, f.a(), f.b()), gy(My, f.a(), f.b());
    for (int i = 0; i < 3; ++i)
        for (int mx = 0; mx < Mx; ++mx)
            for (int mz = 0; mz < Mz; ++mz) {
                ydiff_profile(f, mx, mz, i, g, gy);
                for (int my = 0; my < My; ++my) gradf.cmplx(mx, my, mz, i3j(i, 1)) = gy.get(my);
            }
    f.makeState(sxz, sy);
}

void cross(FlowField& f, FlowField& g, FlowField& fcg, FieldState finalstate) {  // diffops.cpp:2560-2611
    FieldState fxz = f.xzstate(), fy = f.ystate(), gxz = g.xzstate(), gy = g.ystate();
    f.makePhysical();
    g.makePhysical();
    if (!f.geomCongruent(fcg) || fcg.Nd() != 3) fcg.resize(f.Nx(), f.Ny(), f.Nz(), 3, f.Lx(), f.Lz(), f.a(), f.b());
    fcg.setState(Physical, Physical);
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        for (int ny = 0; ny < f.Ny(); ++ny)
            for (int nx = 0; nx < f.Nx(); ++nx)
                for (int nz = 0; nz < f.Nz(); ++nz)
                    fcg.rv(nx, ny, nz, i) = f.rv(nx, ny, nz, j) * g.rv(nx, ny, nz, k) - f.rv(nx, ny, nz, k) * g.rv(nx, ny, nz, j);
    }
    f.makeState(fxz, fy);
    g.makeState(gxz, gy);
    if (finalstate == Spectral) fcg.makeSpectral();
}

void outer(FlowField& f, FlowField& g, FlowField& fg) {  // diffops.cpp:2336-2388
    FieldState fxz = f.xzstate(), fy = f.ystate(), gxz = g.xzstate(), gy = g.ystate();
    f.makePhysical();
    g.makePhysical();
    const int fd = f.Nd(), gd = g.Nd();
    if (!f.geomCongruent(fg) || fg.Nd() != fd * gd)
        fg.resize(f.Nx(), f.Ny(), f.Nz(), fd * gd, f.Lx(), f.Lz(), f.a(), f.b());
    else
        fg.setToZero();
    fg.setState(Physical, Physical);
    for (int i = 0; i < fd; ++i)
        for (int j = 0; j < gd; ++j) {
            const int ij = i * gd + j;
            for (int ny = 0; ny < f.Ny(); ++ny)
                for (int nx = 0; nx < f.Nx(); ++nx)
                    for (int nz = 0; nz < f.Nz(); ++nz) fg.rv(nx, ny, nz, ij) = f.rv(nx, ny, nz, i) * g.rv(nx, ny, nz, j);
        }
    f.makeState(fxz, fy);
    g.makeState(gxz, gy);
    fg.makeSpectral();
}

void div(FlowField& f, FlowField& divf, FieldState finalstate) {  // diffops.cpp:2470-2553 (9d branch)
    FieldState sxz = f.xzstate(), sy = f.ystate();
    f.makeSpectral();
    if (f.Nd() != 9) throw std::runtime_error("div: only 9d fields are on the DNS path");
    if (!f.geomCongruent(divf) || divf.Nd() != 3)
        divf.resize(f.Nx(), f.Ny(), f.Nz(), 3, f.Lx(), f.Lz(), f.a(), f.b());
    else
        divf.setToZero();
    divf.setState(Spectral, Spectral);
    const int My = f.Ny(), Mx = f.Mx(), Mz = f.Mz(), kxmax = f.kxmax(), kzmax = f.kzmax();
    for (int j = 0; j < 3; ++j) {
        for (int my = 0; my < My; ++my)
            for (int mx = 0; mx < Mx; ++mx) {
                int kx = f.kx(mx);
                Complex cxkx(0.0, 2.0 * pi * kx / f.Lx() * zero_last_mode(kx, kxmax, 1));
                for (int mz = 0; mz < Mz; ++mz) {
                    Complex czkz(0.0, 2.0 * pi * mz / f.Lz() * zero_last_mode(mz, kzmax, 1));
                    divf.cmplx(mx, my, mz, j) +=
                        f.cmplx(mx, my, mz, i3j(0, j)) * cxkx + f.cmplx(mx, my, mz, i3j(2, j)) * czkz;
                }
            }
        ComplexChebyCoeff g(My, f.a(), f.b()), gy(My, f.a(), f.b());
        for (int mx = 0; mx < Mx; ++mx)
            for (int mz = 0; mz < Mz; ++mz) {
                ydiff_profile(f, mx, mz, i3j(1, j), g, gy);
                for (int my = 0; my < My; ++my) divf.cmplx(mx, my, mz, j) += gy.get(my);
            }
    }
    f.makeState(sxz, sy);
    if (finalstate == Spectral) divf.makeSpectral();  // diffops.cpp:2554-2556: no Physical branch
}

void rotationalNL(FlowField& u, FlowField& f, FlowField& vort, FieldState finalstate) {  // diffops.cpp:2852-2881
    FieldState uxz = u.xzstate(), uy = u.ystate();
    if (!u.geomCongruent(f) || f.Nd() != 3) f.resize(u.Nx(), u.Ny(), u.Nz(), 3, u.Lx(), u.Lz(), u.a(), u.b());
    f.setState(Physical, Physical);
    if (!u.geomCongruent(vort) || vort.Nd() < 3) vort.resize(u.Nx(), u.Ny(), u.Nz(), 3, u.Lx(), u.Lz(), u.a(), u.b());
    u.makeSpectral();
    curl(u, vort);
    u.makePhysical();
    vort.makePhysical();
    cross(vort, u, f, finalstate);
    if (finalstate == Spectral) f.makeSpectral();
    u.makeState(uxz, uy);
}

void dotgrad(FlowField& u, FlowField& v, FlowField& f, FlowField& grad_v) {  // diffops.cpp:3586-3643
    if (!u.geomCongruent(f) || f.Nd() != 3)
        f.resize(u.Nx(), u.Ny(), u.Nz(), 3, u.Lx(), u.Lz(), u.a(), u.b());
    else
        f.setToZero();
    if (!u.geomCongruent(grad_v) || grad_v.Nd() < 9)
        grad_v.resize(u.Nx(), u.Ny(), u.Nz(), 9, u.Lx(), u.Lz(), u.a(), u.b());
    FieldState uxz = u.xzstate(), uy = u.ystate(), vxz = v.xzstate(), vy = v.ystate();
    v.makeSpectral();
    grad(v, grad_v);
    grad_v.makePhysical();
    u.makePhysical();
    f.setState(Physical, Physical);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int ij = i3j(i, j);
            for (int ny = 0; ny < u.Ny(); ++ny)
                for (int nx = 0; nx < u.Nx(); ++nx)
                    for (int nz = 0; nz < u.Nz(); ++nz) f.rv(nx, ny, nz, i) += u.rv(nx, ny, nz, j) * grad_v.rv(nx, ny, nz, ij);
        }
    f.makeSpectral();
    u.makeState(uxz, uy);
    v.makeState(vxz, vy);
}

void convectionNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState) {  // diffops.cpp:2883-2885
    dotgrad(u, u, f, tmp);
}

void divergenceNL(FlowField& u, FlowField& f, FlowField& uu, FieldState finalstate) {  // diffops.cpp:3110-3134
    FieldState uxz = u.xzstate(), uy = u.ystate();
    if (!u.geomCongruent(f) || f.Nd() != 3) f.resize(u.Nx(), u.Ny(), u.Nz(), 3, u.Lx(), u.Lz(), u.a(), u.b());
    if (!u.geomCongruent(uu) || uu.Nd() < 9) uu.resize(u.Nx(), u.Ny(), u.Nz(), 9, u.Lx(), u.Lz(), u.a(), u.b());
    u.makePhysical();
    outer(u, u, uu);
    uu.makeSpectral();
    div(uu, f, finalstate);
    u.makeState(uxz, uy);
}

void skewsymmetricNL(FlowField& u, FlowField& f, FlowField& tmp, FieldState finalstate) {  // diffops.cpp:3142-3286
    FieldState uxz = u.xzstate(), uy = u.ystate();
    if (!u.geomCongruent(f) || f.Nd() != 3)
        f.resize(u.Nx(), u.Ny(), u.Nz(), 3, u.Lx(), u.Lz(), u.a(), u.b());
    else
        f.setToZero();
    f.setState(Physical, Physical);
    if (!u.geomCongruent(tmp) || tmp.Nd() < 9)
        tmp.resize(u.Nx(), u.Ny(), u.Nz(), 9, u.Lx(), u.Lz(), u.a(), u.b());
    else
        tmp.setToZero();
    u.makeSpectral();
    FlowField& grad_u = tmp;
    grad(u, grad_u);
    grad_u.makePhysical();
    u.makePhysical();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int ij = i3j(i, j);
            for (int ny = 0; ny < u.Ny(); ++ny)
                for (int nx = 0; nx < u.Nx(); ++nx)
                    for (int nz = 0; nz < u.Nz(); ++nz)
                        f.rv(nx, ny, nz, i) += 0.5 * u.rv(nx, ny, nz, j) * grad_u.rv(nx, ny, nz, ij);
        }
    FlowField& uu = tmp;
    outer(u, u, uu);
    uu.makeSpectral();
    f.makeSpectral();
    const int My = u.Ny(), Mx = u.Mx(), Mz = u.Mz(), kxmax = u.kxmax(), kzmax = u.kzmax();
    ComplexChebyCoeff tp(My, u.a(), u.b()), tpy(My, u.a(), u.b());
    for (int i = 0; i < 3; ++i) {
        const int i0 = i3j(i, 0), i1 = i3j(i, 1), i2 = i3j(i, 2);
        for (int my = 0; my < My; ++my)
            for (int mx = 0; mx < Mx; ++mx) {
                int kx = u.kx(mx);
                Complex Dx(0.0, 2 * pi * kx / u.Lx() * zero_last_mode(kx, kxmax, 1));
                for (int mz = 0; mz < Mz; ++mz) {
                    Complex Dz(0.0, 2 * pi * mz / u.Lz() * zero_last_mode(mz, kzmax, 1));
                    f.cmplx(mx, my, mz, i) += 0.5 * (Dx * uu.cmplx(mx, my, mz, i0) + Dz * uu.cmplx(mx, my, mz, i2));
                }
            }
        for (int mx = 0; mx < Mx; ++mx)
            for (int mz = 0; mz < Mz; ++mz) {
                ydiff_profile(uu, mx, mz, i1, tp, tpy);
                for (int my = 0; my < My; ++my) f.cmplx(mx, my, mz, i) += 0.5 * tpy.get(my);
            }
    }
    if (finalstate == Physical) f.makePhysical();
    u.makeState(uxz, uy);
}

void navierstokesNL(FlowField& u, const ChebyCoeff& Ubase, const ChebyCoeff& Wbase, FlowField& f, FlowField& tmp,
                    DNSFlags& flags) {  // nse.cpp:12-91
    FieldState finalstate = Spectral;
    if (flags.rotation != 0.0) finalstate = Physical;
    if (flags.nonlinearity == LinearAboutProfile)
        throw std::runtime_error("LinearAboutProfile is not on the DNS::advance path restated here");
    for (int ny = 0; ny < u.Ny(); ++ny) {
        u.cmplx(0, ny, 0, 0) += Complex(Ubase[ny], 0.0);
        u.cmplx(0, ny, 0, 2) += Complex(Wbase[ny], 0.0);
    }
    u.cmplx(0, 0, 0, 1) -= Complex(flags.Vsuck, 0.);
    switch (flags.nonlinearity) {
        case Rotational: rotationalNL(u, f, tmp, finalstate); break;
        case Convection: convectionNL(u, f, tmp, finalstate); break;
        case SkewSymmetric: skewsymmetricNL(u, f, tmp, finalstate); break;
        case Divergence: divergenceNL(u, f, tmp, finalstate); break;
        case Alternating:
            divergenceNL(u, f, tmp, finalstate);
            flags.nonlinearity = Alternating_;
            break;
        case Alternating_:
            convectionNL(u, f, tmp, finalstate);
            flags.nonlinearity = Alternating;
            break;
        default: throw std::runtime_error("navierstokesNL: unknown method");
    }
    if (flags.rotation != 0.0) {
        u.makePhysical();
        for (int ny = 0; ny < f.Ny(); ++ny)
            for (int nx = 0; nx < f.Nx(); ++nx)
                for (int nz = 0; nz < f.Nz(); ++nz) {
                    f.rv(nx, ny, nz, 0) -= flags.rotation * u.rv(nx, ny, nz, 1);
                    f.rv(nx, ny, nz, 1) += flags.rotation * u.rv(nx, ny, nz, 0);
                }
        u.makeSpectral();
        f.makeSpectral();
    }
    for (int ny = 0; ny < u.Ny(); ++ny) {
        u.cmplx(0, ny, 0, 0) -= Complex(Ubase[ny], 0.0);
        u.cmplx(0, ny, 0, 2) -= Complex(Wbase[ny], 0.0);
    }
    u.cmplx(0, 0, 0, 1) += Complex(flags.Vsuck, 0.);
    u.makeSpectral();
}

ChebyCoeff laminarProfile(Real nu, MeanConstraint constraint, Real dPdx, Real Ubulk, Real Vsuck, Real a, Real b,
                          Real ua, Real ub, int Ny) {  // nse.cpp:823-943
    ChebyCoeff u(Ny, a, b, Spectral);
    Real H = b - a;
    Real mu = Vsuck * H / nu;
    if (std::abs(mu) == 0.0) {
        if (constraint == BulkVelocity) {
            u[0] = 0.125 * (ub + ua) + 0.75 * Ubulk;
            u[1] = 0.5 * (ub - ua);
            u[2] = 0.375 * (ub + ua) - 0.75 * Ubulk;
        } else {
            dPdx *= square((b - a) / 2);
            u[0] = 0.5 * (ub + ua) - 0.25 * dPdx / nu;
            u[1] = 0.5 * (ub - ua);
            u[2] = 0.25 * dPdx / nu;
        }
        return u;
    }
    u.state = Physical;
    Real ub_ua = ub - ua;
    Real expm1_mu = std::expm1(-mu);
    Real dPdx_HH_nu = dPdx * square(H) / nu;
    std::vector<Real> y = chebypoints(Ny, a, b);
    if (std::abs(mu) > 1e-01) {
        if (constraint == BulkVelocity) {
            Real k = -1.0 / expm1_mu - 1 / mu;
            dPdx_HH_nu = mu * (Ubulk - ua - ub_ua * k) / (0.5 - k);
        }
        for (int i = 0; i < Ny; i++) {
            Real yhat = (y[i] - a) / H;
            Real expm1_mu_yhat = std::expm1(-mu * yhat);
            u[i] = ua + (ub_ua)*expm1_mu_yhat / expm1_mu + dPdx_HH_nu * (yhat - expm1_mu_yhat / expm1_mu) / mu;
        }
    } else {
        if (constraint == BulkVelocity) {
            Real mu2 = mu * mu;
            Real A = ub - ua;
            Real B = ua + ub - 2 * Ubulk;
            dPdx_HH_nu = 6 * B + mu * (A + mu * (B / 10 + mu2 * (-B / 1400 + mu2 * B / 126000)));
        }
        const Real one_2 = 1.0 / 2, one_12 = 1.0 / 12, one_24 = 1.0 / 24, one_720 = 1.0 / 720, one_1440 = 1.0 / 1440,
                   one_30240 = 1.0 / 30240, one_120960 = 1.0 / 120960;
        for (int i = 0; i < Ny; i++) {
            Real yh = (y[i] - a) / H, yh2 = yh * yh, yh3 = yh2 * yh, yh4 = yh3 * yh, yh_1 = yh - 1;
            Real taylor1 =
                yh + mu * (-one_2 * yh * yh_1 +
                           mu * (one_12 * yh * yh_1 * (2 * yh - 1) +
                                 mu * (-one_24 * square(yh * yh_1) +
                                       mu * (one_720 * yh * yh_1 * (2 * yh - 1) * (3 * yh2 - 3 * yh - 1) +
                                             mu * (-one_1440 * square(yh * yh_1) * (2 * yh2 - 2 * yh - 1) +
                                                   mu * (one_30240 * yh * yh_1 * (2 * yh - 1) *
                                                         (3 * yh4 - 6 * yh3 + 3 * yh + 1)))))));
            Real taylor2 =
                one_2 * yh * yh_1 + mu * (-one_12 * yh * yh_1 * (2 * yh - 1) +
                                          mu * (one_24 * square(yh * yh_1) +
                                                mu * (-one_720 * yh * yh_1 * (2 * yh - 1) * (3 * yh2 - 3 * yh - 1) +
                                                      mu * (one_1440 * square(yh * yh_1) * (2 * yh2 - 2 * yh - 1) +
                                                            mu * (-one_30240 * yh * yh_1 * (2 * yh - 1) *
                                                                      (3 * yh4 - 6 * yh3 + 3 * yh + 1) +
                                                                  mu * (one_120960 * square(yh * yh_1) *
                                                                        (3 * yh4 - 6 * yh3 + 4 * yh + 2)))))));
            u[i] = ua + ub_ua * taylor1 + dPdx_HH_nu * taylor2;
        }
    }
    u.makeSpectral();
    return u;
}

// ============================================================================
// NSE (nse.cpp)
// ============================================================================
void NSE::setup(const std::vector<FlowField>& fields) {  // nse.cpp:212-289 member initialisers
    const FlowField& u = fields[0];
    if (u.Nd() != 3) throw std::runtime_error("NSE: velocity must have 3 components");
    if (flags_.dealias_y()) throw std::runtime_error("DealiasY/DealiasXYZ are not restated here");
    Nx_ = u.Nx(); Nz_ = u.Nz(); Mx_ = u.Mx(); My_ = u.Ny(); Mz_ = u.Mz();
    Nyd_ = My_;
    kxd_max_ = flags_.dealias_xz() ? u.Nx() / 3 - 1 : u.kxmax();
    kzd_max_ = flags_.dealias_xz() ? u.Nz() / 3 - 1 : u.kzmax();
    Lx_ = u.Lx(); Lz_ = u.Lz(); a_ = u.a(); b_ = u.b();
    kxmax_ = u.kxmax(); kzmax_ = u.kzmax();
    for (auto* c : {&uk_, &vk_, &wk_, &Pk_, &Pyk_, &Ruk_, &Rvk_, &Rwk_}) *c = ComplexChebyCoeff(Nyd_, a_, b_);
    const bool nine = flags_.nonlinearity != Rotational;
    tmp_.resize(u.Nx(), u.Ny(), u.Nz(), nine ? 9 : 3, u.Lx(), u.Lz(), u.a(), u.b());
}

NSE::NSE(const std::vector<FlowField>& fields, const DNSFlags& flags) : flags_(flags) {
    setup(fields);
    createCFBaseFlow();
    initCFConstraint(fields[0]);
}

NSE::NSE(const std::vector<FlowField>& fields, const std::vector<ChebyCoeff>& base, const DNSFlags& flags)
    : flags_(flags) {  // nse.cpp:291-366
    setup(fields);
    Ubase_ = base[0];
    Wbase_ = base[1];
    initCFConstraint(fields[0]);
}

void NSE::createCFBaseFlow() {  // nse.cpp:589-632
    switch (flags_.baseflow) {
        case ZeroBase:
            Ubase_ = ChebyCoeff(My_, a_, b_);
            Wbase_ = ChebyCoeff(My_, a_, b_);
            break;
        case LinearBase:
            Ubase_ = ChebyCoeff(My_, a_, b_);
            Wbase_ = ChebyCoeff(My_, a_, b_);
            Ubase_[1] = 1;
            break;
        case ParabolicBase:
            Ubase_ = ChebyCoeff(My_, a_, b_);
            Wbase_ = ChebyCoeff(My_, a_, b_);
            Ubase_[0] = 0.5;
            Ubase_[2] = -0.5;
            break;
        case SuctionBase:
            Ubase_ = laminarProfile(flags_.nu, PressureGradient, 0, flags_.Ubulk, flags_.Vsuck, a_, b_, -0.5, 0.5, My_);
            Wbase_ = ChebyCoeff(My_, a_, b_);
            break;
        case LaminarBase:
            Ubase_ = laminarProfile(flags_.nu, flags_.constraint, flags_.dPdx, flags_.Ubulk, flags_.Vsuck, a_, b_,
                                    flags_.ulowerwall, flags_.uupperwall, My_);
            Wbase_ = laminarProfile(flags_.nu, flags_.constraint, flags_.dPdz, flags_.Wbulk, flags_.Vsuck, a_, b_,
                                    flags_.wlowerwall, flags_.wupperwall, My_);
            break;
        default: throw std::runtime_error("NSE: ArbitraryBase needs an explicit base flow");
    }
}

static Real profile_mean(const FlowField& u, int i) {  // getUbulk: Re(u.profile(0,0,i)).mean(), diffops.cpp:3853
    ChebyCoeff p(u.Ny(), u.a(), u.b());
    for (int ny = 0; ny < u.Ny(); ++ny) p[ny] = u.cmplx(0, ny, 0, i).real();
    Real m = p.mean();
    if (std::abs(m) < 1e-15) m = 0.0;
    return m;
}

static Real profile_dslope(const FlowField& u, int i, const ChebyCoeff* add) {  // dudy_b - dudy_a, flowfield.cpp:3939
    ChebyCoeff p(u.Ny(), u.a(), u.b()), dp(u.Ny(), u.a(), u.b());
    for (int ny = 0; ny < u.Ny(); ++ny) p[ny] = u.cmplx(0, ny, 0, i).real() + (add ? (*add)[ny] : 0.0);
    diff(p, dp);
    return dp.eval_b() - dp.eval_a();
}

void NSE::initCFConstraint(const FlowField& u) {  // nse.cpp:634-671
    UbulkBase_ = Ubase_.mean();
    ChebyCoeff Ubasey(My_, a_, b_);
    diff(Ubase_, Ubasey);
    Ubaseyy_ = ChebyCoeff(My_, a_, b_);
    diff(Ubasey, Ubaseyy_);
    WbulkBase_ = Wbase_.mean();
    ChebyCoeff Wbasey(My_, a_, b_);
    diff(Wbase_, Wbasey);
    Wbaseyy_ = ChebyCoeff(My_, a_, b_);
    diff(Wbasey, Wbaseyy_);
    UbulkAct_ = UbulkBase_ + profile_mean(u, 0);
    WbulkAct_ = WbulkBase_ + profile_mean(u, 2);
    const Real Ly = b_ - a_;
    dPdxAct_ = flags_.nu * profile_dslope(u, 0, &Ubase_) / Ly;
    dPdzAct_ = flags_.nu * profile_dslope(u, 2, nullptr) / Ly;  // nse.cpp:656-660 adds Wbase to component 0
    if (flags_.constraint == BulkVelocity) {
        UbulkRef_ = flags_.Ubulk;
        WbulkRef_ = flags_.Wbulk;
    } else {
        dPdxAct_ = flags_.dPdx;
        dPdxRef_ = flags_.dPdx;
        dPdzAct_ = flags_.dPdz;
        dPdzRef_ = flags_.dPdz;
    }
}

void NSE::reset_lambda(const std::vector<Real>& lambda_t) {  // nse.cpp:673-705
    lambda_t_ = lambda_t;
    tausolver_.assign(lambda_t.size(), std::vector<TauSolver>((size_t)Mx_ * Mz_));
    const Real c = 4.0 * square(pi) * flags_.nu;
    for (size_t j = 0; j < lambda_t.size(); ++j)
        for (int mx = 0; mx < Mx_; ++mx) {
            int kx = (mx <= Nx_ / 2) ? mx : mx - Nx_;
            for (int mz = 0; mz < Mz_; ++mz) {
                int kz = mz;
                Real lambda = lambda_t[j] + c * (square(kx / Lx_) + square(kz / Lz_));
                if ((kx != kxmax_ || kz != kzmax_) && (!flags_.dealias_xz() || !isAliasedMode(kx, kz)))
                    tausolver_[j][(size_t)mx * Mz_ + mz] =
                        TauSolver(kx, kz, Lx_, Lz_, a_, b_, lambda, flags_.nu, Nyd_, flags_.taucorrection);
            }
        }
}

void NSE::reset_gradp(Real dPdx, Real dPdz) {  // nse.cpp:723-734
    flags_.constraint = PressureGradient;
    flags_.dPdx = dPdx; flags_.dPdz = dPdz; flags_.Ubulk = 0.0; flags_.Wbulk = 0.0;
    dPdxRef_ = dPdx; dPdzRef_ = dPdz; UbulkRef_ = 0.0; WbulkRef_ = 0.0;
}

void NSE::reset_bulkv(Real Ubulk, Real Wbulk) {  // nse.cpp:735-746
    flags_.constraint = BulkVelocity;
    flags_.Ubulk = Ubulk; flags_.Wbulk = Wbulk; flags_.dPdx = 0.0; flags_.dPdz = 0.0;
    UbulkRef_ = Ubulk; WbulkRef_ = Wbulk; dPdxRef_ = 0.0; dPdzRef_ = 0.0;
}

void NSE::nonlinear(std::vector<FlowField>& in, std::vector<FlowField>& out) {  // nse.cpp:383-391
    navierstokesNL(in[0], Ubase_, Wbase_, out[0], tmp_, flags_);
    if (flags_.dealias_xz()) out[0].zeroPaddedModes();
}

void NSE::linear(const std::vector<FlowField>& in, std::vector<FlowField>& out) {  // nse.cpp:393-477
    const FlowField& u = in[0];
    const int kxmax = u.kxmax(), kzmax = u.kzmax();
    for (int mx = 0; mx < Mx_; ++mx) {
        const int kx = u.kx(mx);
        for (int mz = 0; mz < Mz_; ++mz) {
            const int kz = mz;
            if ((kx == kxmax || kz == kzmax) || (flags_.dealias_xz() && isAliasedMode(kx, kz))) break;
            for (int ny = 0; ny < Nyd_; ++ny) {
                uk_.set(ny, flags_.nu * u.cmplx(mx, ny, mz, 0));
                vk_.set(ny, flags_.nu * u.cmplx(mx, ny, mz, 1));
                wk_.set(ny, flags_.nu * u.cmplx(mx, ny, mz, 2));
                Pk_.set(ny, in[1].cmplx(mx, ny, mz, 0));
            }
            diff2(uk_.re, Ruk_.re); diff2(uk_.im, Ruk_.im);
            diff2(vk_.re, Rvk_.re); diff2(vk_.im, Rvk_.im);
            diff2(wk_.re, Rwk_.re); diff2(wk_.im, Rwk_.im);
            diff(Pk_, Pyk_);
            const Real kappa2 = 4 * pi * pi * (square(kx / Lx_) + square(kz / Lz_));
            const Complex Dx(0.0, 2 * pi * kx / Lx_ * zero_last_mode(kx, kxmax, 1));
            const Complex Dz(0.0, 2 * pi * kz / Lz_ * zero_last_mode(kz, kzmax, 1));
            for (int ny = 0; ny < Nyd_; ++ny) {
                out[0].cmplx(mx, ny, mz, 0) = Ruk_.get(ny) - kappa2 * uk_.get(ny) - Dx * Pk_.get(ny);
                out[0].cmplx(mx, ny, mz, 1) = Rvk_.get(ny) - kappa2 * vk_.get(ny) - Pyk_.get(ny);
                out[0].cmplx(mx, ny, mz, 2) = Rwk_.get(ny) - kappa2 * wk_.get(ny) - Dz * Pk_.get(ny);
            }
            if (kx == 0 && kz == 0) {
                for (int ny = 0; ny < My_; ++ny) out[0].cmplx(mx, ny, mz, 0) += Complex(flags_.nu * Ubaseyy_[ny], 0);
                for (int ny = 0; ny < My_; ++ny) out[0].cmplx(mx, ny, mz, 2) += Complex(flags_.nu * Wbaseyy_[ny], 0);
                if (flags_.constraint == PressureGradient) {
                    out[0].cmplx(mx, 0, mz, 0) -= Complex(dPdxRef_, 0);
                    out[0].cmplx(mx, 0, mz, 2) -= Complex(dPdzRef_, 0);
                } else {
                    Real Ly = b_ - a_;
                    diff(uk_, Ruk_);
                    diff(wk_, Rwk_);
                    Real dPdxAct = (Ruk_.re.eval_b() - Ruk_.re.eval_a()) / Ly;
                    Real dPdzAct = (Rwk_.re.eval_b() - Rwk_.re.eval_a()) / Ly;
                    ChebyCoeff Ubasey(My_, a_, b_), Wbasey(My_, a_, b_);
                    diff(Ubase_, Ubasey);
                    diff(Wbase_, Wbasey);
                    if (Ubase_.N() != 0) dPdxAct += flags_.nu * (Ubasey.eval_b() - Ubasey.eval_a()) / Ly;
                    if (Wbase_.N() != 0) dPdzAct += flags_.nu * (Wbasey.eval_b() - Wbasey.eval_a()) / Ly;
                    out[0].cmplx(mx, 0, mz, 0) -= Complex(dPdxAct, 0);
                    out[0].cmplx(mx, 0, mz, 2) -= Complex(dPdzAct, 0);
                }
            }
        }
    }
}

void NSE::solve(std::vector<FlowField>& out, const std::vector<FlowField>& rhs, int s) {  // nse.cpp:479-575
    const int kxmax = out[0].kxmax(), kzmax = out[0].kzmax();
    for (int mx = 0; mx < Mx_; ++mx) {
        const int kx = out[0].kx(mx);
        for (int mz = 0; mz < Mz_; ++mz) {
            const int kz = mz;
            if ((kx == kxmax || kz == kzmax) || (flags_.dealias_xz() && isAliasedMode(kx, kz))) break;
            for (int ny = 0; ny < Nyd_; ++ny) {
                Ruk_.set(ny, rhs[0].cmplx(mx, ny, mz, 0));
                Rvk_.set(ny, rhs[0].cmplx(mx, ny, mz, 1));
                Rwk_.set(ny, rhs[0].cmplx(mx, ny, mz, 2));
            }
            const TauSolver& ts = tausolver_[s][(size_t)mx * Mz_ + mz];
            if (kx != 0 || kz != 0) {
                ts.solve(uk_, vk_, wk_, Pk_, Ruk_, Rvk_, Rwk_);
            } else {
                for (int ny = 0; ny < My_; ++ny) Ruk_.re[ny] += flags_.nu * Ubaseyy_[ny];
                for (int ny = 0; ny < My_; ++ny) Rwk_.re[ny] += flags_.nu * Wbaseyy_[ny];
                if (flags_.constraint == PressureGradient) {
                    Ruk_.re[0] -= dPdxRef_;
                    Rwk_.re[0] -= dPdzRef_;
                    ts.solve(uk_, vk_, wk_, Pk_, Ruk_, Rvk_, Rwk_);
                } else {
                    ts.solve(uk_, vk_, wk_, Pk_, dPdxAct_, dPdzAct_, Ruk_, Rvk_, Rwk_, UbulkRef_ - UbulkBase_,
                             WbulkRef_ - WbulkBase_);
                }
            }
            const bool realmode = (kx == 0 && kz == 0) || (out[0].Nx() % 2 == 0 && kx == kxmax && kz == 0) ||
                                  (out[0].Nz() % 2 == 0 && kz == kzmax && kx == 0) ||
                                  (out[0].Nx() % 2 == 0 && out[0].Nz() % 2 == 0 && kx == kxmax && kz == kzmax);
            for (int ny = 0; ny < Nyd_; ++ny) {
                if (realmode) {
                    out[0].cmplx(mx, ny, mz, 0) = Complex(uk_.re[ny], 0.0);
                    out[0].cmplx(mx, ny, mz, 1) = Complex(vk_.re[ny], 0.0);
                    out[0].cmplx(mx, ny, mz, 2) = Complex(wk_.re[ny], 0.0);
                    out[1].cmplx(mx, ny, mz, 0) = Complex(Pk_.re[ny], 0.0);
                } else {
                    out[0].cmplx(mx, ny, mz, 0) = uk_.get(ny);
                    out[0].cmplx(mx, ny, mz, 1) = vk_.get(ny);
                    out[0].cmplx(mx, ny, mz, 2) = wk_.get(ny);
                    out[1].cmplx(mx, ny, mz, 0) = Pk_.get(ny);
                }
            }
        }
    }
}

// ============================================================================
// Time stepping (dnsalgo.cpp, dns.cpp)
// ============================================================================
DNSAlgorithm::DNSAlgorithm(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags)
    : flags_(flags), numfields_((int)fields.size()), t_(flags.t0), nse_(std::move(nse)) {}

Real DNSAlgorithm::CFL(FlowField& u0) const {  // dnsalgo.cpp:63-67, flowfield.cpp:4035-4068
    FlowField u(u0);
    u.makePhysical();
    ChebyCoeff U = nse_->Ubase(), W = nse_->Wbase();
    U.makePhysical();
    W.makePhysical();
    std::vector<Real> y = chebypoints(u.Ny(), u.a(), u.b());
    Real dx[3] = {u.Lx() / u.Nx(), 0.0, u.Lz() / u.Nz()};
    Real cfl = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int ny = 0; ny < u.Ny(); ++ny) {
            dx[1] = (ny == 0 || ny == u.Ny() - 1) ? y[0] - y[1] : (y[ny - 1] - y[ny + 1]) / 2.0;
            Real Ui = (i == 0) ? U[ny] : (i == 2) ? W[ny] : 0.0;
            for (int nx = 0; nx < u.Nx(); ++nx)
                for (int nz = 0; nz < u.Nz(); ++nz) cfl = std::max(cfl, (u.rv(nx, ny, nz, i) + Ui) / dx[i]);
        }
    cfl *= flags_.dealias_xz() ? 2.0 * pi / 3.0 * flags_.dt : pi * flags_.dt;
    return cfl;
}

MultistepDNS::MultistepDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags)
    : DNSAlgorithm(fields, std::move(nse), flags) {  // dnsalgo.cpp:102-179
    switch (flags.timestepping) {
        case CNFE1:
        case SBDF1: order_ = 1; eta_ = 1.0; alpha_ = {-1.0}; beta_ = {1.0}; break;
        case SBDF2: order_ = 2; eta_ = 1.5; alpha_ = {-2.0, 0.5}; beta_ = {2.0, -1.0}; break;
        case SBDF3: order_ = 3; eta_ = 11.0 / 6.0; alpha_ = {-3.0, 1.5, -1.0 / 3.0}; beta_ = {3.0, -3.0, 1.0}; break;
        case SBDF4:
            order_ = 4; eta_ = 25.0 / 12.0;
            alpha_ = {-4.0, 3.0, -4.0 / 3.0, 0.25};
            beta_ = {4.0, -6.0, 4.0, -1.0};
            break;
        default: throw std::runtime_error("MultistepDNS: not a multistep method");
    }
    lambda_t_ = {eta_ / flags_.dt};
    nse_->reset_lambda(lambda_t_);
    std::vector<FlowField> tmpar(fields);
    for (auto& f : tmpar) f.setToZero();
    fields_.assign(order_, tmpar);
    nonlf_.assign(order_, tmpar);
    Ninitsteps_ = order_ - 1;
    countdown_ = Ninitsteps_;
}

void MultistepDNS::reset_dt(Real dt) {  // dnsalgo.cpp:191-200
    flags_.dt = dt;
    lambda_t_ = {eta_ / flags_.dt};
    nse_->reset_lambda(lambda_t_);
    countdown_ = Ninitsteps_;
}

void MultistepDNS::advance(std::vector<FlowField>& fieldsn, int Nsteps) {  // dnsalgo.cpp:202-253
    const int J = order_ - 1;
    std::vector<FlowField> rhs{fieldsn[0]};
    const int len = (int)rhs.size();
    fields_[0] = fieldsn;
    for (int step = 0; step < Nsteps; ++step) {
        nse_->nonlinear(fields_[0], nonlf_[0]);
        for (int l = 0; l < len; ++l) {
            rhs[l].setToZero();
            for (int j = 0; j < order_; ++j) {
                const Real a = -alpha_[j] / flags_.dt;
                const Real b = -beta_[j];
                rhs[l].add(a, fields_[j][l], b, nonlf_[j][l]);
            }
        }
        nse_->solve(fields_[J], rhs);
        for (int j = order_ - 1; j > 0; --j)
            for (int l = 0; l < numfields_; ++l) {
                std::swap(nonlf_[j][l], nonlf_[j - 1][l]);
                std::swap(fields_[j][l], fields_[j - 1][l]);
            }
        t_ += flags_.dt;
    }
    fieldsn = fields_[0];
}

bool MultistepDNS::push(std::vector<FlowField>& fields) {  // dnsalgo.cpp:287-315
    for (int j = order_ - 1; j > 0; --j)
        for (int l = 0; l < numfields_; ++l) {
            std::swap(nonlf_[j][l], nonlf_[j - 1][l]);
            std::swap(fields_[j][l], fields_[j - 1][l]);
        }
    if (order_ > 1) {
        fields_[1] = fields;
        nse_->nonlinear(fields_[1], nonlf_[1]);
    }
    t_ += flags_.dt;
    --countdown_;
    return full();
}

CNABstyleDNS::CNABstyleDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags)
    : DNSAlgorithm(fields, std::move(nse), flags), fj1_(fields), fj_(fields) {  // dnsalgo.cpp:510-576
    switch (flags_.timestepping) {
        case CNAB2:
            order_ = 2; Nsubsteps_ = 1; Ninitsteps_ = 1; full_ = false;
            alpha_ = {0.5}; beta_ = {0.5}; gamma_ = {1.5}; zeta_ = {-0.5};
            break;
        case SMRK2:
            order_ = 2; Nsubsteps_ = 3; Ninitsteps_ = 0; full_ = true;
            alpha_ = {29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0};
            beta_ = {37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0};
            gamma_ = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
            zeta_ = {0.0, -17.0 / 60.0, -5.0 / 12.0};
            break;
        default: throw std::runtime_error("CNABstyleDNS: not a CNAB-style method");
    }
    for (auto& f : fj1_) f.setToZero();
    for (auto& f : fj_) f.setToZero();
    lambda_t_.resize(Nsubsteps_);
    for (int j = 0; j < Nsubsteps_; ++j) lambda_t_[j] = 1.0 / (beta_[j] * flags_.dt);
    nse_->reset_lambda(lambda_t_);
}

void CNABstyleDNS::reset_dt(Real dt) {  // dnsalgo.cpp:584-596
    flags_.dt = dt;
    for (int j = 0; j < Nsubsteps_; ++j) lambda_t_[j] = 1.0 / (beta_[j] * flags_.dt);
    nse_->reset_lambda(lambda_t_);
    if (flags_.timestepping == CNAB2) full_ = false;
}

bool CNABstyleDNS::push(std::vector<FlowField>& fields) {  // dnsalgo.cpp:598-607
    for (int l = 0; l < numfields_; ++l) std::swap(fj_[l], fj1_[l]);
    nse_->nonlinear(fields, fj_);  // const_cast in the reference: the caller's u is transformed in place
    t_ += flags_.dt;
    full_ = true;
    return full_;
}

void CNABstyleDNS::advance(std::vector<FlowField>& fields, int Nsteps) {  // dnsalgo.cpp:641-697
    std::vector<FlowField> rhs{fields[0]};
    std::vector<FlowField> lt(rhs);
    const int len = (int)rhs.size();
    for (int n = 0; n < Nsteps; ++n) {
        for (int j = 0; j < Nsubsteps_; ++j) {
            for (int l = 0; l < numfields_; ++l) std::swap(fj_[l], fj1_[l]);
            nse_->nonlinear(fields, fj_);
            Real a_b = alpha_[j] / beta_[j];
            Real g_b = gamma_[j] / beta_[j];
            Real z_b = zeta_[j] / beta_[j];
            nse_->linear(fields, lt);
            for (int l = 0; l < len; ++l) {
                rhs[l] = fields[l];
                rhs[l].scale(lambda_t_[j]);
                rhs[l].add(a_b, lt[l]);
                rhs[l].add(-g_b, fj_[l], -z_b, fj1_[l]);
            }
            nse_->solve(fields, rhs, j);
        }
        t_ += flags_.dt;
    }
}

RungeKuttaDNS::RungeKuttaDNS(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse, const DNSFlags& flags)
    : DNSAlgorithm(fields, std::move(nse), flags), Qj1_(fields), Qj_(fields) {  // dnsalgo.cpp:351-395
    if (flags_.timestepping != CNRK2) throw std::runtime_error("RungeKuttaDNS: only CNRK2");
    order_ = 2; Nsubsteps_ = 3; Ninitsteps_ = 0;
    A_ = {0.0, -5.0 / 9.0, -153.0 / 128.0};
    B_ = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    C_ = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
    for (auto& f : Qj1_) f.setToZero();
    for (auto& f : Qj_) f.setToZero();
    lambda_t_.resize(Nsubsteps_);
    for (int j = 0; j < Nsubsteps_; ++j) lambda_t_[j] = 1.0 / (C_[j] * flags_.dt);
    nse_->reset_lambda(lambda_t_);
}

void RungeKuttaDNS::reset_dt(Real dt) {  // dnsalgo.cpp:404-414
    flags_.dt = dt;
    for (int j = 0; j < Nsubsteps_; ++j) lambda_t_[j] = 1.0 / (C_[j] * flags_.dt);
    nse_->reset_lambda(lambda_t_);
}

void RungeKuttaDNS::advance(std::vector<FlowField>& fields, int Nsteps) {  // dnsalgo.cpp:416-459
    std::vector<FlowField> rhs{fields[0]};
    std::vector<FlowField> lt(rhs);
    const int len = (int)rhs.size();
    for (int n = 0; n < Nsteps; ++n) {
        for (int j = 0; j < Nsubsteps_; ++j) {
            Qj_[0].scale(A_[j]);
            nse_->nonlinear(fields, Qj1_);
            Qj_[0].add(-1.0, Qj1_[0]);
            nse_->linear(fields, lt);
            const Real B_C = B_[j] / C_[j];
            for (int l = 0; l < len; ++l) {
                rhs[l] = lt[l];
                rhs[l].add(lambda_t_[j], fields[l], B_C, Qj_[l]);
            }
            nse_->solve(fields, rhs, j);
        }
        t_ += flags_.dt;
    }
}

std::unique_ptr<DNSAlgorithm> DNS::newAlgorithm(const std::vector<FlowField>& fields, std::shared_ptr<NSE> nse,
                                                const DNSFlags& flags) {  // dns.cpp:95-115
    switch (flags.timestepping) {
        case CNFE1:
        case SBDF1:
        case SBDF2:
        case SBDF3:
        case SBDF4: return std::make_unique<MultistepDNS>(fields, nse, flags);
        case CNRK2: return std::make_unique<RungeKuttaDNS>(fields, nse, flags);
        case SMRK2:
        case CNAB2: return std::make_unique<CNABstyleDNS>(fields, nse, flags);
    }
    throw std::runtime_error("DNS::newAlgorithm: unknown algorithm");
}

DNS::DNS(const std::vector<FlowField>& fields, const DNSFlags& flags) {  // dns.cpp:22-39
    main_nse_ = std::make_shared<NSE>(fields, flags);
    main_ = newAlgorithm(fields, main_nse_, flags);
    if (!main_->full() && flags.initstepping != flags.timestepping) {
        DNSFlags initflags = flags;
        initflags.timestepping = flags.initstepping;
        init_nse_ = std::make_shared<NSE>(fields, initflags);
        init_ = newAlgorithm(fields, init_nse_, initflags);
    }
}

DNS::DNS(const std::vector<FlowField>& fields, const std::vector<ChebyCoeff>& base, const DNSFlags& flags_) {
    DNSFlags flags = flags_;  // dns.cpp:41-62
    flags.baseflow = ArbitraryBase;
    main_nse_ = std::make_shared<NSE>(fields, base, flags);
    main_ = newAlgorithm(fields, main_nse_, flags);
    if (!main_->full() && flags.initstepping != flags.timestepping) {
        DNSFlags initflags = flags;
        initflags.timestepping = flags.initstepping;
        init_nse_ = std::make_shared<NSE>(fields, base, initflags);
        init_ = newAlgorithm(fields, init_nse_, initflags);
    }
}

void DNS::advance(std::vector<FlowField>& fields, int Nsteps) {  // dns.cpp:135-163
    if (!main_->full() && !init_) throw std::runtime_error("DNS::advance: uninitialised multistep algorithm");
    for (size_t j = 1; j < fields.size(); ++j)
        if (!fields[j].geomCongruent(fields[0]))
            fields[j].resize(fields[0].Nx(), fields[0].Ny(), fields[0].Nz(), fields[j].Nd(), fields[0].Lx(),
                             fields[0].Lz(), fields[0].a(), fields[0].b());
    int n = 0;
    while (!main_->full() && n < Nsteps) {
        main_->push(fields);
        init_->advance(fields, 1);
        ++n;
    }
    main_->advance(fields, Nsteps - n);
}

void DNS::reset_dt(Real dt) {  // dns.cpp:180-185
    main_->reset_dt(dt);
    if (init_) init_->reset_dt(dt);
}

void DNS::reset_gradp(Real dPdx, Real dPdz) {  // dns.cpp:221-226
    if (init_nse_) init_nse_->reset_gradp(dPdx, dPdz);
    main_nse_->reset_gradp(dPdx, dPdz);
}

void DNS::reset_bulkv(Real Ubulk, Real Wbulk) {  // dns.cpp:227-232
    if (init_nse_) init_nse_->reset_bulkv(Ubulk, Wbulk);
    main_nse_->reset_bulkv(Ubulk, Wbulk);
}

// ============================================================================
// PressureSolver (poissonsolver.cpp:270-431)
// ============================================================================
// p solves lapl p = -div nonl(u + U) with p = 0 at the walls, mode by mode
// (PoissonSolver::solve, :146-171: a HelmholtzSolver of lambda = 4 pi^2
// (kx^2/Lx^2 + kz^2/Lz^2), nu = 1, per mode; the divergence as div's 3d branch,
// diffops.cpp:2488-2516). Every mode but (0,0) then gets g = c exp(-mu (y-a)) +
// d exp(mu (y-b)), mu^2 = lambda, chosen so that dp/dy = nu v'' at y = a, b
// (:356-427); g is evaluated at the Chebyshev points and transformed.
void pressure_solve(const FlowField& u0, const ChebyCoeff& Ubase, const ChebyCoeff& Wbase, Real nu, Real Vsuck,
                    NonlinearMethod nonl, FlowField& p) {
    const int Nx = u0.Nx(), Ny = u0.Ny(), Nz = u0.Nz(), Mz = u0.Mz();
    const Real Lx = u0.Lx(), Lz = u0.Lz(), a = u0.a(), b = u0.b();
    FlowField u(u0);
    FlowField f(Nx, Ny, Nz, 3, Lx, Lz, a, b), tmp;
    DNSFlags fl;
    fl.nonlinearity = nonl;
    fl.Vsuck = Vsuck;
    fl.nu = nu;
    navierstokesNL(u, Ubase, Wbase, f, tmp, fl);
    p = FlowField(Nx, Ny, Nz, 1, Lx, Lz, a, b);
    const int kxmax = f.kxmax(), kzmax = f.kzmax();
    const std::vector<Real> y = chebypoints(Ny, a, b);
    const Real H = b - a;
    ChebyCoeff gr(Ny, a, b), gi(Ny, a, b), gyr(Ny, a, b), gyi(Ny, a, b);
    for (int mx = 0; mx < Nx; ++mx)
        for (int mz = 0; mz < Mz; ++mz) {
            const int kx = f.kx(mx), kz = mz;
            const Complex cx(0.0, 2.0 * pi * kx / Lx * zero_last_mode(kx, kxmax, 1));
            const Complex cz(0.0, 2.0 * pi * kz / Lz * zero_last_mode(kz, kzmax, 1));
            // -div f
            for (int my = 0; my < Ny; ++my) {
                gr[my] = f.cmplx(mx, my, mz, 1).real();
                gi[my] = f.cmplx(mx, my, mz, 1).imag();
            }
            diff(gr, gyr);
            diff(gi, gyi);
            ChebyCoeff fr(Ny, a, b), fi(Ny, a, b);
            for (int my = 0; my < Ny; ++my) {
                Complex d = f.cmplx(mx, my, mz, 0) * cx + f.cmplx(mx, my, mz, 2) * cz;
                d += Complex(gyr[my], gyi[my]);
                d *= -1.0;
                fr[my] = d.real();
                fi[my] = d.imag();
            }
            const Real lambda = 4.0 * pi * pi * ((kx / Lx) * (kx / Lx) + (kz / Lz) * (kz / Lz));
            const HelmholtzSolver h(Ny, a, b, lambda);
            ChebyCoeff pr(Ny, a, b), pi_(Ny, a, b);
            h.solve(pr, fr, 0.0, 0.0);
            h.solve(pi_, fi, 0.0, 0.0);
            if (mx != 0 || mz != 0) {  // dp/dy = nu v_yy at the walls
                ChebyCoeff pyr(Ny, a, b), pyi(Ny, a, b), vr(Ny, a, b), vi(Ny, a, b), vyyr(Ny, a, b), vyyi(Ny, a, b);
                diff(pr, pyr);
                diff(pi_, pyi);
                for (int my = 0; my < Ny; ++my) {
                    vr[my] = u.cmplx(mx, my, mz, 1).real();
                    vi[my] = u.cmplx(mx, my, mz, 1).imag();
                }
                diff2(vr, vyyr);
                diff2(vi, vyyi);
                const Complex alpha = nu * Complex(vyyr.eval_a(), vyyi.eval_a()) - Complex(pyr.eval_a(), pyi.eval_a());
                const Complex beta = nu * Complex(vyyr.eval_b(), vyyi.eval_b()) - Complex(pyr.eval_b(), pyi.eval_b());
                const Real mu = std::sqrt(lambda);
                const Real delta = mu * (1 - std::exp(-2 * mu * H));
                const Complex c = (-alpha + beta * std::exp(-mu * H)) / delta;
                const Complex d = (beta - alpha * std::exp(-mu * H)) / delta;
                ChebyCoeff g_r(Ny, a, b, Physical), g_i(Ny, a, b, Physical);
                for (int my = 0; my < Ny; ++my) {
                    const Complex g = c * std::exp(-mu * (y[my] - a)) + d * std::exp(mu * (y[my] - b));
                    g_r[my] = g.real();
                    g_i[my] = g.imag();
                }
                g_r.makeSpectral();
                g_i.makeSpectral();
                for (int my = 0; my < Ny; ++my) {
                    pr[my] += g_r[my];
                    pi_[my] += g_i[my];
                }
            }
            for (int my = 0; my < Ny; ++my) p.cmplx(mx, my, mz, 0) = Complex(pr[my], pi_[my]);
        }
}

}  // namespace cfref
