/* cfgpu.hpp -- C++ host facade of the cfgpu C ABI with channelflow's DNS
 * signatures, so that a channelflow program swaps its DNS object and keeps the
 * rest of its code.
 *
 *   chflow::DNS dns(fields, flags);         ->  cfgpu::DNS<FlowField, DNSFlags> dns(fields, flags);
 *   dns.advance(fields, n);                     dns.advance(fields, n);
 *   dns.reset_dt(dt); dns.CFL(fields[0]);       (same calls)
 *
 * Header-only over include/cfgpu.h (link libcfgpu.so). The field and flag
 * types are template parameters, so the header needs none of channelflow's
 * headers; they are used through the members channelflow gives them:
 *   FieldT (channelflow/flowfield.h:52-345): Nx(), Ny(), Nz(), Nd(), Lx(), Lz(),
 *     a(), b(), makeSpectral(), cmplx(mx, my, mz, i) -> std::complex<double>&,
 *     contiguous in channelflow's serial spectral layout from cmplx(0,0,0,0)
 *     (flowfield.h:374-401), and copy construction.
 *   FlagsT (channelflow/dnsflags.h:81-145): the public fields baseflow,
 *     constraint, timestepping, initstepping, nonlinearity, dealiasing,
 *     taucorrection, nu, Vsuck, rotation, dPdx, dPdz, Ubulk, Wbulk,
 *     ulowerwall, uupperwall, wlowerwall, wupperwall, t0, dt, with the enum
 *     values of dnsflags.h:26-38.
 *   ChebyT (channelflow/chebyshev.h, ChebyCoeff): ChebyT(N, a, b) and
 *     operator[](int) -> double&.
 *
 * Errors: channelflow reports them through cferror(), which prints and exits
 * (cfbasics/cfbasics.h:272-278). The library never exits; this facade throws
 * cfgpu::Error (status code and cfg_last_error() text). A program that wants
 * channelflow's behaviour catches it at the top and calls cferror.
 *
 * Threading is that of the C ABI: one host thread at a time.
 */
#ifndef CFGPU_HPP
#define CFGPU_HPP

#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfgpu.h"

namespace cfgpu {

typedef double Real;

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline void check(int rc, const char* call) {
    if (rc != CFG_OK) throw Error(rc, std::string(call) + ": " + cfg_last_error());
}

/* Chebyshev coefficients of a base flow: the default ChebyT. */
class Cheby {
   public:
    Cheby() = default;
    Cheby(int N, Real a, Real b) : c_(N, 0.0), a_(a), b_(b) {}
    int N() const { return (int)c_.size(); }
    int numModes() const { return N(); }
    Real a() const { return a_; }
    Real b() const { return b_; }
    Real& operator[](int n) { return c_[n]; }
    const Real& operator[](int n) const { return c_[n]; }

   private:
    std::vector<Real> c_;
    Real a_ = -1, b_ = 1;
};

/* DNSFlags -> cfg_flags (field meanings of channelflow/dnsflags.h:106-140). */
template <class FlagsT>
cfg_flags to_cfg(const FlagsT& f) {
    cfg_flags c;
    std::memset(&c, 0, sizeof c);
    c.baseflow = static_cast<int>(f.baseflow);
    c.constraint = static_cast<int>(f.constraint);
    c.timestepping = static_cast<int>(f.timestepping);
    c.initstepping = static_cast<int>(f.initstepping);
    c.nonlinearity = static_cast<int>(f.nonlinearity);
    c.dealiasing = static_cast<int>(f.dealiasing);
    c.taucorrection = f.taucorrection ? 1 : 0;
    c.nu = f.nu;
    c.Vsuck = f.Vsuck;
    c.rotation = f.rotation;
    c.dPdx = f.dPdx;
    c.dPdz = f.dPdz;
    c.Ubulk = f.Ubulk;
    c.Wbulk = f.Wbulk;
    c.ulowerwall = f.ulowerwall;
    c.uupperwall = f.uupperwall;
    c.wlowerwall = f.wlowerwall;
    c.wupperwall = f.wupperwall;
    c.t0 = f.t0;
    c.dt = f.dt;
    return c;
}

template <class FieldT>
cfg_grid grid_of(const FieldT& u) {
    cfg_grid g;
    std::memset(&g, 0, sizeof g);
    g.Nx = u.Nx();
    g.Ny = u.Ny();
    g.Nz = u.Nz();
    g.Lx = u.Lx();
    g.Lz = u.Lz();
    g.a = u.a();
    g.b = u.b();
    return g;
}

/* First double of a spectral field's buffer (serial layout). */
template <class FieldT>
double* spectral_data(FieldT& f) {
    return reinterpret_cast<double*>(&f.cmplx(0, 0, 0, 0));
}
template <class FieldT>
const double* spectral_data(const FieldT& f) {
    return reinterpret_cast<const double*>(&const_cast<FieldT&>(f).cmplx(0, 0, 0, 0));
}

/* Spectral copies of fields[0] (u, 3 components) and fields[1] (q), as
 * DNS::DNS copies the caller's fields (dns.cpp:22-62); q may be absent. */
template <class FieldT>
struct SpectralPair {
    std::vector<FieldT> f;
    explicit SpectralPair(const std::vector<FieldT>& fields) : f(fields) {
        if (f.empty()) throw Error(CFG_ERR_ARG, "cfgpu::DNS: no velocity field");
        if (f[0].Nd() != 3) throw Error(CFG_ERR_ARG, "cfgpu::DNS: fields[0] must have 3 components");
        for (auto& x : f) x.makeSpectral();
    }
    const double* u() const { return spectral_data(f[0]); }
    const double* q() const { return f.size() > 1 ? spectral_data(f[1]) : nullptr; }
};

/* channelflow's DNS (channelflow/dns.h:30-97) on the GPU: DNS::advance and the
 * calls its callers make around it. The state stays on the device between
 * advance() calls; the caller's fields are uploaded only when they differ from
 * what the previous advance() returned. */
template <class FieldT, class FlagsT, class ChebyT = Cheby>
class DNS {
   public:
    /* DNS(fields, flags) -- dns.h:34 */
    DNS(const std::vector<FieldT>& fields, const FlagsT& flags) : flags_(flags) {
        create(fields, nullptr, nullptr);
    }
    /* DNS(fields, {Ubase, Wbase}, flags) -- dns.h:35 */
    DNS(const std::vector<FieldT>& fields, const std::vector<ChebyT>& base, const FlagsT& flags) : flags_(flags) {
        if (base.size() != 2) throw Error(CFG_ERR_ARG, "cfgpu::DNS: base must hold Ubase and Wbase");
        const int Ny = fields.at(0).Ny();
        std::vector<Real> U(Ny), W(Ny);
        for (int n = 0; n < Ny; ++n) {
            U[n] = base[0][n];
            W[n] = base[1][n];
        }
        create(fields, U.data(), W.data());
    }
    DNS(const DNS&) = delete;
    DNS& operator=(const DNS&) = delete;
    ~DNS() {
        if (h_) (void)cfg_dns_destroy(h_);
    }

    /* DNS::advance(fields, nSteps) -- dns.h:43, dns.cpp:135-163: fields[0] = u
     * and fields[1] = q in and out, spectral on return. */
    void advance(std::vector<FieldT>& fields, int nSteps = 1) {
        if (fields.size() < 2) throw Error(CFG_ERR_ARG, "cfgpu::DNS::advance: fields must hold u and q");
        FieldT& u = fields[0];
        FieldT& q = fields[1];
        u.makeSpectral();
        q.makeSpectral();
        double* ud = spectral_data(u);
        double* qd = spectral_data(q);
        if (!same(ud, last_u_) || !same(qd, last_q_)) check(cfg_dns_set_state(h_, ud, qd), "cfg_dns_set_state");
        check(cfg_dns_advance(h_, nSteps), "cfg_dns_advance");
        check(cfg_dns_get_state(h_, ud, qd), "cfg_dns_get_state");
        last_u_.assign(ud, ud + nu_);
        last_q_.assign(qd, qd + nq_);
    }

    /* dns.h:54-57 */
    void reset_dt(Real dt) {
        check(cfg_dns_reset_dt(h_, dt), "cfg_dns_reset_dt");
        flags_.dt = dt;
    }
    /* hipGraph replay of repeating steps: off by default, bit-identical to eager */
    void set_graphs(bool on) { check(cfg_dns_set_graphs(h_, on ? 1 : 0), "cfg_dns_set_graphs"); }
    void reset_time(Real t) { check(cfg_dns_reset_time(h_, t), "cfg_dns_reset_time"); }
    void reset_gradp(Real dPdx, Real dPdz) { check(cfg_dns_reset_gradp(h_, dPdx, dPdz), "cfg_dns_reset_gradp"); }
    void reset_bulkv(Real Ubulk, Real Wbulk) { check(cfg_dns_reset_bulkv(h_, Ubulk, Wbulk), "cfg_dns_reset_bulkv"); }

    /* dns.h:65-81. CFL(u) is that of the DNS's own velocity, which is the
     * caller's fields[0] after advance() (the reference's callers pass that
     * field, programs/simulateflow.cpp:140-142); a u edited since is uploaded
     * first. */
    Real nu() const { return flags_.nu; }
    Real dt() const { return flags_.dt; }
    Real CFL(FieldT& u) {
        u.makeSpectral();
        double* ud = spectral_data(u);
        if (!same(ud, last_u_)) {
            check(cfg_dns_set_state(h_, ud, nullptr), "cfg_dns_set_state");
            last_u_.assign(ud, ud + nu_);
        }
        double c = 0;
        check(cfg_dns_cfl(h_, &c), "cfg_dns_cfl");
        return c;
    }
    Real time() const {
        double t = 0;
        check(cfg_dns_time(h_, &t), "cfg_dns_time");
        return t;
    }
    Real dPdx() const {
        double v = 0;
        check(cfg_dns_dpdx(h_, &v), "cfg_dns_dpdx");
        return v;
    }
    Real Ubulk() const {
        double v = 0;
        check(cfg_dns_ubulk(h_, &v), "cfg_dns_ubulk");
        return v;
    }
    const ChebyT& Ubase() const { return Ubase_; }
    const ChebyT& Wbase() const { return Wbase_; }
    const FlagsT& flags() const { return flags_; }

    cfg_dns* handle() { return h_; }

   private:
    void create(const std::vector<FieldT>& fields, const double* U, const double* W) {
        const SpectralPair<FieldT> s(fields);
        const cfg_grid g = grid_of(s.f[0]);
        const cfg_flags c = to_cfg(flags_);
        nu_ = (size_t)2 * 3 * g.Nx * g.Ny * (g.Nz / 2 + 1);
        nq_ = nu_ / 3;
        check(cfg_dns_create(&g, &c, s.u(), s.q(), U, W, &h_), "cfg_dns_create");
        std::vector<double> Ub(g.Ny), Wb(g.Ny);
        check(cfg_dns_base(h_, Ub.data(), Wb.data()), "cfg_dns_base");
        Ubase_ = ChebyT(g.Ny, g.a, g.b);
        Wbase_ = ChebyT(g.Ny, g.a, g.b);
        for (int n = 0; n < g.Ny; ++n) {
            Ubase_[n] = Ub[n];
            Wbase_[n] = Wb[n];
        }
        // the device now holds the caller's fields; q absent means zero
        last_u_.assign(s.u(), s.u() + nu_);
        if (s.q()) last_q_.assign(s.q(), s.q() + nq_);
        else last_q_.assign(nq_, 0.0);
    }
    static bool same(const double* p, const std::vector<double>& v) {
        return !v.empty() && std::memcmp(p, v.data(), v.size() * sizeof(double)) == 0;
    }
    cfg_dns* h_ = nullptr;
    FlagsT flags_;
    ChebyT Ubase_, Wbase_;
    size_t nu_ = 0, nq_ = 0;
    std::vector<double> last_u_, last_q_;
};

/* Many independent integrations in one object: the serial loops of
 * programs/edgetracking.cpp:472-530 (bisection members),
 * channelflow/multiShootingDSI.cpp:70-86 (shooting segments) and
 * channelflow/cfdsi.cpp:678-775 (Jacobian probes) hand their members to one
 * batch. Each member behaves as a DNS of its own. */
template <class FieldT, class FlagsT>
class DNSBatch {
   public:
    DNSBatch() { check(cfg_batch_create(&h_), "cfg_batch_create"); }
    DNSBatch(const DNSBatch&) = delete;
    DNSBatch& operator=(const DNSBatch&) = delete;
    ~DNSBatch() {
        if (h_) (void)cfg_batch_destroy(h_);
    }
    /* a member from fields = {u, q} as DNS(fields, flags); returns its id */
    int add(const std::vector<FieldT>& fields, const FlagsT& flags) {
        const SpectralPair<FieldT> s(fields);
        const cfg_grid g = grid_of(s.f[0]);
        const cfg_flags c = to_cfg(flags);
        int id = -1;
        check(cfg_batch_add(h_, &g, &c, s.u(), s.q(), &id), "cfg_batch_add");
        return id;
    }
    void remove(int id) { check(cfg_batch_remove(h_, id), "cfg_batch_remove"); }
    /* every member nSteps steps */
    void advance(int nSteps) { check(cfg_batch_advance(h_, nSteps), "cfg_batch_advance"); }
    /* the listed members their own step counts, together */
    void advance(const std::vector<int>& ids, const std::vector<int>& nSteps) {
        if (ids.size() != nSteps.size()) throw Error(CFG_ERR_ARG, "cfgpu::DNSBatch::advance: size mismatch");
        check(cfg_batch_advance_members(h_, (int)ids.size(), ids.data(), nSteps.data()), "cfg_batch_advance_members");
    }
    /* member state into fields = {u, q} (spectral) */
    void fetch(int id, std::vector<FieldT>& fields) {
        if (fields.size() < 2) throw Error(CFG_ERR_ARG, "cfgpu::DNSBatch::fetch: fields must hold u and q");
        fields[0].makeSpectral();
        fields[1].makeSpectral();
        check(cfg_batch_get_state(h_, id, spectral_data(fields[0]), spectral_data(fields[1])), "cfg_batch_get_state");
    }
    void store(int id, std::vector<FieldT>& fields) {
        fields[0].makeSpectral();
        fields[1].makeSpectral();
        check(cfg_batch_set_state(h_, id, spectral_data(fields[0]), spectral_data(fields[1])), "cfg_batch_set_state");
    }
    void reset_dt(int id, Real dt) { check(cfg_batch_reset_dt(h_, id, dt), "cfg_batch_reset_dt"); }
    /* the member's velocity was finite after its last advance (an Error with
       code CFG_ERR_NONFINITE from advance names the members for which not) */
    bool finite(int id) {
        int f = 0;
        check(cfg_batch_finite(h_, id, &f), "cfg_batch_finite");
        return f != 0;
    }
    Real time(int id) {
        double t = 0;
        check(cfg_batch_time(h_, id, &t), "cfg_batch_time");
        return t;
    }
    Real CFL(int id) {
        double c = 0;
        check(cfg_batch_cfl(h_, id, &c), "cfg_batch_cfl");
        return c;
    }
    Real dPdx(int id) {
        double v = 0;
        check(cfg_batch_dpdx(h_, id, &v), "cfg_batch_dpdx");
        return v;
    }
    cfg_batch* handle() { return h_; }

   private:
    cfg_batch* h_ = nullptr;
};

}  // namespace cfgpu

#endif /* CFGPU_HPP */
