// Kernel launchers of the MI355X DNS time step (implemented in cfg_kernels.hip).
// All buffers are device pointers; all arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cfg {

// Per-kernel timing with HIP events on the launching stream (off by default).
enum KernelId { K_YINV = 0, K_XINV, K_ZPASS, K_XFWD, K_YFWD, K_COMBINE, K_TAU, K_LINEAR, K_SETUP, K_COUNT };

struct KTimer {
    bool on = false;
    hipStream_t st = nullptr;
    struct Rec {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double total_ms[K_COUNT] = {};
    long count[K_COUNT] = {};

    hipEvent_t take() {
        hipEvent_t e;
        if (!pool.empty()) {
            e = pool.back();
            pool.pop_back();
        } else {
            (void)hipEventCreate(&e);
        }
        return e;
    }
    hipEvent_t begin() {
        if (!on) return nullptr;
        hipEvent_t a = take();
        (void)hipEventRecord(a, st);
        return a;
    }
    void end(int id, hipEvent_t a) {
        if (!on || !a) return;
        hipEvent_t b = take();
        (void)hipEventRecord(b, st);
        recs.push_back(Rec{id, a, b});
        if (recs.size() > 4096) flush();
    }
    void flush() {  // waits for the recorded work and accumulates elapsed times
        if (recs.empty()) return;
        (void)hipEventSynchronize(recs.back().b);
        for (auto& r : recs) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            total_ms[r.id] += ms;
            count[r.id] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    void reset() {
        flush();
        for (int i = 0; i < K_COUNT; ++i) {
            total_ms[i] = 0.0;
            count[i] = 0;
        }
    }
    ~KTimer() {
        for (auto& r : recs) {
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Compact spectral layout of one field with nd components:
//   complex element (c, ny, ikz, ikx) at index ((c*Ny + ny)*nkz + ikz)*nkx + ikx,
//   kx = ikx for ikx <= kxd, ikx - nkx otherwise; kz = ikz.
// Only live modes |kx| <= kxd, kz <= kzd are stored (M = nkx*nkz per plane).
struct Geom {
    int Nx, Ny, Nz;
    int kxd, kzd, nkx, nkz, M;
    double Lx, Lz, a, b;
};

struct FFTPlan {
    int N;
    int npass;
    int radix[12];
    const double2* tw;  // exp(-2 pi i j / N), j < N
};

enum NLMethod { NL_ROT = 0, NL_CONV = 1, NL_DIV = 2, NL_SKEW = 3 };

struct YMats {            // y-direction operators, row-major [row][col], Ny x Ny
    const double* T;      // Chebyshev coefficients -> values at y_n
    const double* TD;     // T * (d/dy)
    const double* Ti;     // values -> coefficients
    const double* DTi;    // (d/dy) * Ti
};

// Per-mode tau data for one lambda_t, mode-major: array k (TAU_NA of them:
// pband pup pinvd vband vup vinvd Pp Pm vp vm P0 v0) of mode m is the Ny-vector
// at coef + (m*TAU_NA + k)*Ny; scalars i00 i01 i10 i11 sigma0_Nb1 sigma0_Nb at
// scal + m*6.
constexpr int TAU_NA = 12;
struct TauSet {
    double* coef;
    double* scal;
    double lambda_t;
};

struct SolveConst {  // kx=kz=0 mode data (nse.cpp:505-547)
    const double* nuUbaseyy;  // [Ny] nu * U''(coefficients)
    const double* nuWbaseyy;
    double dPdxRef, dPdzRef;
    double umean, wmean;  // UbulkRef - UbulkBase, WbulkRef - WbulkBase
    int bulk;             // 1: BulkVelocity constraint
    int reserved;         // 0 (launch tables are compared byte-wise: no padding)
    double* dPdx_out;     // device scalars written by the bulk solve
    double* dPdz_out;
};

struct LinearConst {  // kx=kz=0 constants of NSE::linear (nse.cpp:437-472)
    const double* nuUbaseyy;
    const double* nuWbaseyy;
    double dPdxRef, dPdzRef;
    int bulk;
    int reserved;  // 0
    double nuUbase_slope, nuWbase_slope;  // nu (U'(b) - U'(a)) / Ly etc.
};

// ---- per-member launch tables ----------------------------------------------
// Every kernel of the time step runs over n members of one grid class at once
// (blockIdx.y or blockIdx.z = member). Pointers and parameters that differ
// between members come from a device-resident table of these structs.

constexpr int MAX_TERMS = 8;

// Tau solve of one member: rhs = sum_j coef[j] in[j] (3 components, compact
// layout), solved into u (3 components) and q (1 component).
struct TauMember {
    const double* in[MAX_TERMS];
    double coef[MAX_TERMS];
    int nterms;
    int taucorr;
    double nu;
    TauSet set;
    SolveConst sc;
    double* u;
    double* q;
};

// NSE::linear of one member: out = nu (u'' - kappa^2 u) - grad q (+ constants).
struct LinearMember {
    const double* u;
    const double* q;
    double* out;
    double nu;
    LinearConst c;
};

// Nonlinear term of one member: f = N(u + base flow).
struct NLMember {
    const double* u;     // compact velocity [3][Ny][M] complex
    double* f;           // compact nonlinear term [3][Ny][M] complex
    const double* base;  // [4][Ny]: U(y), dU/dy(y), W(y), dW/dy(y) at the y_n
    double Vsuck, rotation;
};

// ---- tau ------------------------------------------------------------------
void launch_tau_setup(const Geom& g, const TauSet& s, double nu, bool taucorr, hipStream_t st);
// PressureSolver::solve (poissonsolver.cpp:326-431) of one field: q from the
// nonlinear term f of u + U and from v; the pressure factors come from set.
struct PressArgs {
    const double* f;  // compact 3-component nonlinear term
    const double* u;  // compact velocity (v is used)
    double* q;        // compact pressure out
    TauSet set;
    double nu;
};
void launch_pressure(const Geom& g, const YMats& ym, const double* hb, const PressArgs& pa, hipStream_t st);

// hb: rows of the Helmholtz systems' operator B, [3][Ny] (Geometry::HB)
void launch_tau_solve(const Geom& g, const double* hb, int n, const TauMember* tab, hipStream_t st);

// ---- elementwise ----------------------------------------------------------
// out = sum_j coef[j] * in[j], over n doubles (up to 8 terms)
void launch_combine(long n, int nterms, const double* coef, const double* const* in, double* out, hipStream_t st);
// flags[m] = 1 (left unchanged otherwise) if fields[m][0..n) holds a NaN or an infinity
void launch_nonfinite(int nfields, const double* const* fields, long n, double* flags, hipStream_t st);
void launch_linear(const Geom& g, int n, const LinearMember* tab, hipStream_t st);

// ---- nonlinear term ---------------------------------------------------------
// LDS bytes a block may use (the device's opt-in limit); default 64 KiB.
void set_lds_budget(size_t bytes);

// Work buffers of member 0; member p uses base + p * (per-member size below).
struct NLWork {
    double* yp;  // [6][Ny][M] complex
    double* xp;  // [9][Ny][Nx][nkz] complex
    double* zo;  // [9][Ny][Nx][nkz] complex
    double* fo;  // [9][Ny][M] complex
};
inline long nlwork_y_doubles(const Geom& g) { return 2L * 9 * g.Ny * g.M; }         // yp and fo
inline long nlwork_x_doubles(const Geom& g) { return 2L * 9 * g.Ny * g.Nx * g.nkz; }  // xp and zo

// Tile shapes of the nonlinear-term launches (0: the default rule).
struct NLTune {
    int zxb = 0;  // (y, x) lines per block of the LDS z-pass
    int xl = 0;   // kz lines per block of the LDS x-FFTs
    int zwb = 0;      // wavefronts per block of the unstaged wavefront z-pass (0: 4)
    int zstage = -1;  // wavefront z-pass stages its inputs in LDS: 1 / 0, or -1 the grid's rule
    int xwb = 0;  // wavefronts per block of the wavefront x-FFTs (0: 4)
    int xwbf = 0; // the same for the forward wavefront x-FFT (0: xwb)
};

void launch_nonlinear(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int method, int n,
                      const NLMember* tab, bool rotating, const NLWork& w, const NLTune& tn, hipStream_t st,
                      KTimer* kt = nullptr);

// Times the candidate tilings on these work buffers (contents overwritten)
// and returns the fastest. Results are identical for every tiling.
// Largest (u_i + U_i)/dx_i over the grid points of each of n fields, into
// out[n] (device): FlowField::CFLfactor(Ubase, Wbase), flowfield.cpp:4035-4068.
// Uses the work buffers of the nonlinear term.
void launch_cfl(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int n, const NLMember* tab,
                const NLWork& w, double* out, hipStream_t st);

NLTune tune_nonlinear(const Geom& g, const YMats& ym, const FFTPlan& px, const FFTPlan& pz, int method, int n,
                      const NLMember* tab, bool rotating, const NLWork& w, hipStream_t st);

}  // namespace cfg
