// Tick form of channelflow's time stepping (dns.cpp, dnsalgo.cpp, nse.cpp)
// and the batched tick executor.
#include <cstring>

#include "cfg_step.hpp"

namespace cfg {

// ---------------------------------------------------------------------------
// NSE
// ---------------------------------------------------------------------------
NSECtx::NSECtx(Exec& ex, std::shared_ptr<FlowModel> model, const cfg_flags& f)
    : ex_(ex), model_(std::move(model)), flags_(f), method_(f.nonlinearity) {
    dP_ = PBuf(ex_.pool, 2);
    dPdx_ = model_->dPdxAct0();
    dPdz_ = model_->dPdzAct0();
    const double h[2] = {dPdx_, dPdz_};
    HIPCHK(hipMemcpyAsync(dP_.p, h, sizeof(h), hipMemcpyHostToDevice, ex_.st));
    HIPCHK(hipStreamSynchronize(ex_.st));
}

void NSECtx::reset_lambda(const std::vector<double>& lambda_t) {  // nse.cpp:673-705
    tau_.clear();
    for (double l : lambda_t)
        tau_.push_back(tau_for(*ex_.geo, flags_.nu, l, flags_.taucorrection != 0, ex_.st, ex_.pool, &ex_.kt));
}

void NSECtx::nonlinear(const double* u, double* f, TickOps& ops) {  // nse.cpp:383-391
    int m;
    switch (method_) {
        case CFG_Rotational: m = NL_ROT; break;
        case CFG_Convection: m = NL_CONV; break;
        case CFG_Divergence: m = NL_DIV; break;
        case CFG_SkewSymmetric: m = NL_SKEW; break;
        case CFG_Alternating: m = NL_DIV; method_ = CFG_Alternating_; break;
        default: m = NL_CONV; method_ = CFG_Alternating; break;
    }
    ops.nlon = true;
    ops.method = m;
    ops.nl = NLMember{u, f, model_->base_values(), flags_.Vsuck, flags_.rotation};
}

void NSECtx::linear(const double* u, const double* q, double* out, TickOps& ops) {  // nse.cpp:393-477
    ops.lin = true;
    ops.linm = LinearMember{u, q, out, flags_.nu, model_->linear_const()};
}

void NSECtx::solve(double* u, double* q, std::initializer_list<std::pair<double, const double*>> terms, int s,
                   TickOps& ops) {  // nse.cpp:479-575 with the right-hand side of dnsalgo.cpp
    TauMember& t = ops.tau;
    t = TauMember{};
    int n = 0;
    for (auto& tm : terms) {
        t.coef[n] = tm.first;
        t.in[n] = tm.second;
        ++n;
    }
    t.nterms = n;
    t.taucorr = flags_.taucorrection != 0;
    t.nu = flags_.nu;
    t.set = tau_.at(s)->set;
    t.sc = model_->solve_const(dP_.p);
    t.u = u;
    t.q = q;
    ops.solve = true;
    if (model_->flags().constraint == CFG_BulkVelocity) dP_dirty_ = true;
}

void NSECtx::fetch_dP() {
    if (model_->flags().constraint == CFG_BulkVelocity && dP_dirty_) {
        double h[2];
        HIPCHK(hipMemcpyAsync(h, dP_.p, sizeof(h), hipMemcpyDeviceToHost, ex_.st));
        HIPCHK(hipStreamSynchronize(ex_.st));
        dPdx_ = h[0];
        dPdz_ = h[1];
        dP_dirty_ = false;
    }
}
double NSECtx::dPdx() {
    fetch_dP();
    return dPdx_;
}
double NSECtx::dPdz() {
    fetch_dP();
    return dPdz_;
}

// ---------------------------------------------------------------------------
// Steppers (dnsalgo.cpp)
// ---------------------------------------------------------------------------
class MultistepT : public Stepper {  // dnsalgo.cpp:102-315
   public:
    MultistepT(Exec& ex, NSECtx& nse, const cfg_flags& f) : Stepper(nse, f) {
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
        nse_.reset_lambda({eta_ / flags_.dt});
        const Geom& g = ex.geo->g;
        for (int j = 0; j < order_; ++j) {
            if (order_ > 1) {
                fields_.emplace_back(ex.pool, g);
                fields_.back().u.zero(ex.st);
                fields_.back().q.zero(ex.st);
            }
            nonlf_.emplace_back(ex.pool, ex.nfield);
            nonlf_.back().zero(ex.st);
        }
        countdown_ = order_ - 1;
    }
    bool full() const override { return countdown_ == 0; }
    int ticks_per_step() const override { return 1; }
    int Ninitsteps() const override { return order_ - 1; }
    FieldPair* front() override { return order_ > 1 ? &fields_[0] : nullptr; }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        nse_.reset_lambda({eta_ / flags_.dt});
        countdown_ = order_ - 1;
    }
    void push(StateRef& cur, TickOps& ops) override {  // dnsalgo.cpp:287-315
        rotate();
        if (order_ > 1) {
            FieldPair& f1 = fields_[1];
            ops.add_copy(f1.u.p, cur.u, f1.u.n);
            ops.add_copy(f1.q.p, cur.q, f1.q.n);
            nse_.nonlinear(f1.u.p, nonlf_[1].p, ops);
        }
        t_ += flags_.dt;
        --countdown_;
    }
    // One step (dnsalgo.cpp:210-235). For order > 1 the state is the history
    // front; order 1 steps the state in place.
    void tick(int, StateRef& cur, TickOps& ops) override {
        const int J = order_ - 1;
        const double dt = flags_.dt;
        double* su = order_ > 1 ? fields_[0].u.p : cur.u;
        nse_.nonlinear(su, nonlf_[0].p, ops);
        if (order_ == 1) {
            nse_.solve(cur.u, cur.q, {{-alpha_[0] / dt, su}, {-beta_[0], nonlf_[0].p}}, 0, ops);
        } else if (order_ == 2) {
            nse_.solve(fields_[J].u.p, fields_[J].q.p,
                       {{-alpha_[0] / dt, su}, {-beta_[0], nonlf_[0].p},
                        {-alpha_[1] / dt, fields_[1].u.p}, {-beta_[1], nonlf_[1].p}}, 0, ops);
        } else if (order_ == 3) {
            nse_.solve(fields_[J].u.p, fields_[J].q.p,
                       {{-alpha_[0] / dt, su}, {-beta_[0], nonlf_[0].p},
                        {-alpha_[1] / dt, fields_[1].u.p}, {-beta_[1], nonlf_[1].p},
                        {-alpha_[2] / dt, fields_[2].u.p}, {-beta_[2], nonlf_[2].p}}, 0, ops);
        } else {
            nse_.solve(fields_[J].u.p, fields_[J].q.p,
                       {{-alpha_[0] / dt, su}, {-beta_[0], nonlf_[0].p},
                        {-alpha_[1] / dt, fields_[1].u.p}, {-beta_[1], nonlf_[1].p},
                        {-alpha_[2] / dt, fields_[2].u.p}, {-beta_[2], nonlf_[2].p},
                        {-alpha_[3] / dt, fields_[3].u.p}, {-beta_[3], nonlf_[3].p}}, 0, ops);
        }
        rotate();
        t_ += dt;
    }

   private:
    void rotate() {
        for (int j = order_ - 1; j > 0; --j) {
            std::swap(nonlf_[j], nonlf_[j - 1]);
            if (order_ > 1) {
                std::swap(fields_[j].u, fields_[j - 1].u);
                std::swap(fields_[j].q, fields_[j - 1].q);
            }
        }
    }
    int order_ = 0, countdown_ = 0;
    double eta_ = 0;
    std::vector<double> alpha_, beta_;
    std::vector<FieldPair> fields_;
    std::vector<PBuf> nonlf_;
};

class CNABstyleT : public Stepper {  // dnsalgo.cpp:510-697
   public:
    CNABstyleT(Exec& ex, NSECtx& nse, const cfg_flags& f) : Stepper(nse, f) {
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
        for (auto* b : {&fj_, &fj1_, &lt_}) {
            *b = PBuf(ex.pool, ex.nfield);
            b->zero(ex.st);
        }
        set_lambda();
    }
    bool full() const override { return full_; }
    int ticks_per_step() const override { return nsub_; }
    int Ninitsteps() const override { return ninit_; }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        set_lambda();
        if (flags_.timestepping == CFG_CNAB2) full_ = false;
    }
    void push(StateRef& cur, TickOps& ops) override {
        std::swap(fj_, fj1_);
        nse_.nonlinear(cur.u, fj_.p, ops);
        t_ += flags_.dt;
        full_ = true;
    }
    void tick(int j, StateRef& cur, TickOps& ops) override {  // dnsalgo.cpp:654-690
        std::swap(fj_, fj1_);
        nse_.nonlinear(cur.u, fj_.p, ops);
        const double a_b = alpha_[j] / beta_[j], g_b = gamma_[j] / beta_[j], z_b = zeta_[j] / beta_[j];
        nse_.linear(cur.u, cur.q, lt_.p, ops);
        nse_.solve(cur.u, cur.q, {{lambda_[j], cur.u}, {a_b, lt_.p}, {-g_b, fj_.p}, {-z_b, fj1_.p}}, j, ops);
        if (j == nsub_ - 1) t_ += flags_.dt;
    }

   private:
    void set_lambda() {
        lambda_.resize(nsub_);
        for (int j = 0; j < nsub_; ++j) lambda_[j] = 1.0 / (beta_[j] * flags_.dt);
        nse_.reset_lambda(lambda_);
    }
    int nsub_ = 0, ninit_ = 0;
    bool full_ = false;
    std::vector<double> alpha_, beta_, gamma_, zeta_, lambda_;
    PBuf fj_, fj1_, lt_;
};

class RungeKuttaT : public Stepper {  // dnsalgo.cpp:351-459
   public:
    RungeKuttaT(Exec& ex, NSECtx& nse, const cfg_flags& f) : Stepper(nse, f) {
        A_ = {0.0, -5.0 / 9.0, -153.0 / 128.0};
        B_ = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
        C_ = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
        for (auto* b : {&Qj_, &Qj1_, &lt_}) {
            *b = PBuf(ex.pool, ex.nfield);
            b->zero(ex.st);
        }
        set_lambda();
    }
    bool full() const override { return true; }
    int ticks_per_step() const override { return 3; }
    int Ninitsteps() const override { return 0; }
    void reset_dt(double dt) override {
        flags_.dt = dt;
        set_lambda();
    }
    void tick(int j, StateRef& cur, TickOps& ops) override {  // dnsalgo.cpp:420-446
        nse_.nonlinear(cur.u, Qj1_.p, ops);
        ops.comb = true;
        ops.cn = 2;
        ops.ccoef[0] = A_[j];
        ops.cin[0] = Qj_.p;
        ops.ccoef[1] = -1.0;
        ops.cin[1] = Qj1_.p;
        ops.cout = Qj_.p;
        nse_.linear(cur.u, cur.q, lt_.p, ops);
        nse_.solve(cur.u, cur.q, {{1.0, lt_.p}, {lambda_[j], cur.u}, {B_[j] / C_[j], Qj_.p}}, j, ops);
        if (j == 2) t_ += flags_.dt;
    }

   private:
    void set_lambda() {
        lambda_.resize(3);
        for (int j = 0; j < 3; ++j) lambda_[j] = 1.0 / (C_[j] * flags_.dt);
        nse_.reset_lambda(lambda_);
    }
    std::vector<double> A_, B_, C_, lambda_;
    PBuf Qj_, Qj1_, lt_;
};

std::unique_ptr<Stepper> make_stepper(Exec& ex, NSECtx& nse, const cfg_flags& f) {
    switch (f.timestepping) {
        case CFG_CNFE1:
        case CFG_SBDF1:
        case CFG_SBDF2:
        case CFG_SBDF3:
        case CFG_SBDF4: return std::make_unique<MultistepT>(ex, nse, f);
        case CFG_CNRK2: return std::make_unique<RungeKuttaT>(ex, nse, f);
        default: return std::make_unique<CNABstyleT>(ex, nse, f);
    }
}

// ---------------------------------------------------------------------------
// Integrator (dns.cpp)
// ---------------------------------------------------------------------------
Integrator::Integrator(Exec& ex, const cfg_flags& flags, const double* u_spec, const double* q_spec,
                       const double* Ubase, const double* Wbase)
    : ex_(ex), flags_(flags) {
    check_flags(flags);
    const Geom& g = ex.geo->g;
    own_ = FieldPair(ex.pool, g);
    cur_ = StateRef{own_.u.p, own_.q.p};
    std::vector<double> up = upload_packed(g, 3, u_spec, own_.u.p, ex.st);
    if (q_spec) (void)upload_packed(g, 1, q_spec, own_.q.p, ex.st);
    else own_.q.zero(ex.st);
    cfg_flags f = flags;
    if (Ubase) f.baseflow = CFG_ArbitraryBase;
    flags_ = f;
    model_ = std::make_shared<FlowModel>(*ex.geo, f, up, Ubase, Wbase);
    main_nse_ = std::make_unique<NSECtx>(ex, model_, f);
    main_ = make_stepper(ex, *main_nse_, f);
    if (!main_->full() && f.initstepping != f.timestepping) {  // dns.cpp:26-38
        cfg_flags fi = f;
        fi.timestepping = f.initstepping;
        init_nse_ = std::make_unique<NSECtx>(ex, model_, fi);
        init_ = make_stepper(ex, *init_nse_, fi);
        if (init_->Ninitsteps() != 0) throw Error(CFG_ERR_ARG, "initstepping method needs initialisation itself");
    }
    HIPCHK(hipStreamSynchronize(ex.st));
}

void Integrator::own_state(TickOps* ops) {
    if (cur_.u == own_.u.p) return;
    // the state lives in the multistep history front: copy it out
    if (ops) {
        ops->add_copy(own_.u.p, cur_.u, own_.u.n);
        ops->add_copy(own_.q.p, cur_.q, own_.q.n);
    } else {
        HIPCHK(hipMemcpyAsync(own_.u.p, cur_.u, own_.u.n * sizeof(double), hipMemcpyDeviceToDevice, ex_.st));
        HIPCHK(hipMemcpyAsync(own_.q.p, cur_.q, own_.q.n * sizeof(double), hipMemcpyDeviceToDevice, ex_.st));
    }
    cur_ = StateRef{own_.u.p, own_.q.p};
}

void Integrator::next(TickOps& ops) {  // dns.cpp:135-163, one tick at a time
    ops = TickOps{};
    if (init_sub_ >= 0) {  // the init step that follows a push runs to its end, even once the push filled the history
        init_->tick(init_sub_, cur_, ops);
        if (++init_sub_ == init_->ticks_per_step()) {
            init_sub_ = -1;
            ++steps_;
        }
        return;
    }
    if (!main_->full()) {
        if (!init_) throw Error(CFG_ERR_STATE, "uninitialised multistep algorithm");
        own_state(&ops);
        main_->push(cur_, ops);
        init_sub_ = 0;
        return;
    }
    if (FieldPair* fr = main_->front()) {  // multistep, order > 1: the state is the history front
        if (cur_.u != fr->u.p) {
            ops.add_copy(fr->u.p, cur_.u, fr->u.n);
            ops.add_copy(fr->q.p, cur_.q, fr->q.n);
        }
        main_->tick(0, cur_, ops);
        fr = main_->front();
        cur_ = StateRef{fr->u.p, fr->q.p};
        ++steps_;
        return;
    }
    main_->tick(main_sub_, cur_, ops);
    if (++main_sub_ == main_->ticks_per_step()) {
        main_sub_ = 0;
        ++steps_;
    }
}

void Integrator::set_state(const double* u_spec, const double* q_spec) {
    const Geom& g = ex_.geo->g;
    own_state(nullptr);
    if (u_spec) (void)upload_packed(g, 3, u_spec, own_.u.p, ex_.st);
    if (q_spec) (void)upload_packed(g, 1, q_spec, own_.q.p, ex_.st);
}

void Integrator::get_state(double* u_spec, double* q_spec) {
    const Geom& g = ex_.geo->g;
    if (u_spec) download_packed(g, 3, cur_.u, ex_.nfield, u_spec, ex_.st);
    if (q_spec) download_packed(g, 1, cur_.q, ex_.nscal, q_spec, ex_.st);
}

void Integrator::reset_dt(double dt) {  // dns.cpp:180-185
    main_->reset_dt(dt);
    if (init_) init_->reset_dt(dt);
    flags_.dt = dt;
    init_sub_ = -1;
    main_sub_ = 0;
}

NLMember Integrator::cfl_member() const {
    return NLMember{cur_.u, nullptr, model_->base_values(), flags_.Vsuck, flags_.rotation};
}

double Integrator::cfl_scale() const {
    const double pi = 3.14159265358979323846;
    return flags_.dealiasing == CFG_DealiasXZ ? 2.0 * pi / 3.0 * flags_.dt : pi * flags_.dt;
}

void Integrator::reset_time(double t) {
    main_->reset_time(t);
    if (init_) init_->reset_time(t);
}

void Integrator::reset_gradp(double dPdx, double dPdz) { model_->reset_gradp(dPdx, dPdz); }
void Integrator::reset_bulkv(double Ubulk, double Wbulk) { model_->reset_bulkv(Ubulk, Wbulk); }

// ---------------------------------------------------------------------------
// Exec
// ---------------------------------------------------------------------------
Exec::Exec(const cfg_grid& grid, bool dealias) {
    init_device_limits();
    geo = geometry_for(grid, dealias);
    HIPCHK(hipStreamCreate(&st));
    kt.st = st;
    const Geom& g = geo->g;
    nfield = 2L * 3 * g.Ny * g.M;
    nscal = 2L * g.Ny * g.M;
}

Exec::~Exec() {
    if (st) {
        (void)hipStreamSynchronize(st);
        drop_graphs();
        kt.reset();
        (void)hipStreamDestroy(st);
    }
}

size_t Exec::work_bytes(const Geom& g, int cap) {
    if (cap <= 0) return 0;
    return 2 * device_bytes_taken((size_t)cap * nlwork_y_doubles(g) * sizeof(double)) +
           2 * device_bytes_taken((size_t)cap * nlwork_x_doubles(g) * sizeof(double));
}

size_t Exec::held_bytes() const { return pool.held() + work_bytes(geo->g, cap_); }

void Exec::trim() {
    HIPCHK(hipStreamSynchronize(st));
    drop_graphs();
    pool.trim();
    yp_.alloc(0);
    fo_.alloc(0);
    xp_.alloc(0);
    zo_.alloc(0);
    cap_ = 0;
    ++buf_gen_;
}

void Exec::reserve(int n) {
    if (n <= cap_) return;
    HIPCHK(hipStreamSynchronize(st));  // the old work buffers may still be in use
    const int c = grown_cap(cap_, n);
    const Geom& g = geo->g;
    yp_.alloc((size_t)c * nlwork_y_doubles(g));
    fo_.alloc((size_t)c * nlwork_y_doubles(g));
    xp_.alloc((size_t)c * nlwork_x_doubles(g));
    zo_.alloc((size_t)c * nlwork_x_doubles(g));
    cap_ = c;
    ++buf_gen_;
}

// sequences (seen or captured) kept per class
static constexpr size_t kMaxGraphKeys = 256;

void Exec::set_graphs(bool on) {
    sync();
    graphs_on_ = on;
    if (!on) drop_graphs();
}

void Exec::drop_graphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
    seen_.clear();
}

std::vector<double> Exec::cfl(const std::vector<NLMember>& fields) {
    const int n = (int)fields.size();
    std::vector<double> out(n, 0.0);
    if (n == 0) return out;
    reserve(n);
    if (cflout_.n < (size_t)n) cflout_.alloc(n);
    const NLMember* tab = cache.get(fields, st);
    const NLWork w{yp_.p, xp_.p, zo_.p, fo_.p};
    launch_cfl(geo->g, geo->ymats(), geo->fx.plan, geo->fz.plan, n, tab, w, cflout_.p, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out.data(), cflout_.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return out;
}

std::vector<bool> Exec::nonfinite(const std::vector<const double*>& fields, long n) {
    const int k = (int)fields.size();
    std::vector<bool> out(k, false);
    if (k == 0) return out;
    if (nfout_.n < (size_t)k) nfout_.alloc(k);
    HIPCHK(hipMemsetAsync(nfout_.p, 0, k * sizeof(double), st));
    const double* const* tab = cache.get(fields, st);
    launch_nonfinite(k, tab, n, nfout_.p, st);
    HIPCHK(hipGetLastError());
    std::vector<double> h(k);
    HIPCHK(hipMemcpyAsync(h.data(), nfout_.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < k; ++i) out[i] = h[i] != 0.0;
    return out;
}

// Resolves a tick into launch records: member tables are made resident (a
// copy only for bytes not seen before) and a nonlinear form's tiling is tuned
// on its first launch at a member-count bucket. Nothing here is captured.
//
// The tables of one tick must belong to one generation of the table arena: a
// recycle in the middle of a plan (an arena filled by a later table) would
// leave the earlier records pointing at addresses the later tables overwrite.
// The plan is therefore redone once in the fresh arena, which holds a whole
// tick's tables.
void Exec::plan(const std::vector<TickOps*>& ops) {
    for (int attempt = 0;; ++attempt) {
        const unsigned long gen0 = cache.generation();
        plan_once(ops);
        if (cache.generation() == gen0) return;
        if (attempt > 0) throw Error(CFG_ERR_ARG, "the launch tables of one tick do not fit the table arena");
    }
}

void Exec::plan_once(const std::vector<TickOps*>& ops) {
    const Geom& g = geo->g;
    recs_.clear();
    auto rec = [&](LaunchRec::Kind k) -> LaunchRec& {
        recs_.emplace_back();
        LaunchRec& r = recs_.back();
        std::memset(&r, 0, sizeof r);  // padding too: the record bytes are the graph key
        r.kind = k;
        return r;
    };
    for (TickOps* o : ops)
        for (int i = 0; i < o->ncopy; ++i) {
            LaunchRec& r = rec(LaunchRec::COPY);
            r.dst = o->copy[i].dst;
            r.src = o->copy[i].src;
            r.bytes = o->copy[i].n * sizeof(double);
        }
    const YMats ym = geo->ymats();
    const NLWork w{yp_.p, xp_.p, zo_.p, fo_.p};
    for (int m : {NL_ROT, NL_CONV, NL_DIV, NL_SKEW}) {
        nlt_.clear();
        bool rotating = false;
        for (TickOps* o : ops)
            if (o->nlon && o->method == m) {
                nlt_.push_back(o->nl);
                rotating = rotating || o->nl.rotation != 0.0;
            }
        if (nlt_.empty()) continue;
        const int n = (int)nlt_.size();
        const NLMember* tab = cache.get(nlt_, st);
        int nb = 1;
        while (nb < n) nb *= 2;
        auto tk = tune_.find({m, nb});
        if (tk == tune_.end())  // first launch of this form at this batch size
            tk = tune_.emplace(std::make_pair(m, nb),
                               tune_nonlinear(g, ym, geo->fx.plan, geo->fz.plan, m, n, tab, rotating, w, st)).first;
        LaunchRec& r = rec(LaunchRec::NONLINEAR);
        r.method = m;
        r.n = n;
        r.tab = tab;
        r.rotating = rotating;
        r.tn = tk->second;
    }
    for (TickOps* o : ops)
        if (o->comb) {
            LaunchRec& r = rec(LaunchRec::COMBINE);
            r.cn = o->cn;
            for (int j = 0; j < o->cn; ++j) {
                r.ccoef[j] = o->ccoef[j];
                r.cin[j] = o->cin[j];
            }
            r.cout = o->cout;
        }
    lt_.clear();
    for (TickOps* o : ops)
        if (o->lin) {
            lt_.push_back(o->linm);
        }
    if (!lt_.empty()) {
        LaunchRec& r = rec(LaunchRec::LINEAR);
        r.n = (int)lt_.size();
        r.tab = cache.get(lt_, st);
    }
    tt_.clear();
    for (TickOps* o : ops)
        if (o->solve) tt_.push_back(o->tau);
    if (!tt_.empty()) {
        LaunchRec& r = rec(LaunchRec::TAU);
        r.n = (int)tt_.size();
        r.tab = cache.get(tt_, st);
    }
}

void Exec::issue() {
    const Geom& g = geo->g;
    const YMats ym = geo->ymats();
    const NLWork w{yp_.p, xp_.p, zo_.p, fo_.p};
    for (const LaunchRec& r : recs_) {
        switch (r.kind) {
            case LaunchRec::COPY:
                HIPCHK(hipMemcpyAsync(r.dst, r.src, r.bytes, hipMemcpyDeviceToDevice, st));
                break;
            case LaunchRec::NONLINEAR:
                launch_nonlinear(g, ym, geo->fx.plan, geo->fz.plan, r.method, r.n,
                                 static_cast<const NLMember*>(r.tab), r.rotating != 0, w, r.tn, st, &kt);
                break;
            case LaunchRec::COMBINE: {
                hipEvent_t ev = kt.begin();
                launch_combine(nfield, r.cn, r.ccoef, r.cin, r.cout, st);
                kt.end(K_COMBINE, ev);
                break;
            }
            case LaunchRec::LINEAR: {
                hipEvent_t ev = kt.begin();
                launch_linear(g, r.n, static_cast<const LinearMember*>(r.tab), st);
                kt.end(K_LINEAR, ev);
                break;
            }
            case LaunchRec::TAU: {
                hipEvent_t ev = kt.begin();
                launch_tau_solve(g, geo->HB.p, r.n, static_cast<const TauMember*>(r.tab), st);
                kt.end(K_TAU, ev);
                break;
            }
        }
    }
    HIPCHK(hipGetLastError());
}

void Exec::prepare(const std::vector<TickOps*>& ops) {
    reserve((int)ops.size());
    plan(ops);
}

void Exec::append_key(std::string& key) const {
    key.append(reinterpret_cast<const char*>(recs_.data()), recs_.size() * sizeof(LaunchRec));
}

void Exec::run(const std::vector<TickOps*>& ops) {
    if (ops.empty()) return;
    prepare(ops);
    if (!graphs_on_ || kt.on) {  // timing needs the events between launches
        issue();
        return;
    }
    const unsigned long gen = generation();
    if (gen != graph_gen_) {
        HIPCHK(hipStreamSynchronize(st));  // replays in flight may hold the old graphs
        drop_graphs();
        graph_gen_ = gen;
    }
    std::string key(reinterpret_cast<const char*>(recs_.data()), recs_.size() * sizeof(LaunchRec));
    auto it = graphs_.find(key);
    if (it != graphs_.end()) {
        HIPCHK(hipGraphLaunch(it->second, st));
        ++replays_;
        return;
    }
    if (!seen_.count(key)) {  // first occurrence: eager
        if (seen_.size() + graphs_.size() >= kMaxGraphKeys) {  // membership churn: start over
            HIPCHK(hipStreamSynchronize(st));
            drop_graphs();
        }
        seen_.insert(std::move(key));
        issue();
        return;
    }
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        issue();
    } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
    }
    HIPCHK(hipStreamEndCapture(st, &gr));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCHK(e);
    graphs_.emplace(key, ge);
    seen_.erase(key);
    ++captures_;
    HIPCHK(hipGraphLaunch(ge, st));
}

}  // namespace cfg
