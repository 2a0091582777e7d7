// Time stepping of channelflow's DNS (dns.cpp, dnsalgo.cpp, nse.cpp) as a
// sequence of "ticks". A tick is one nonlinear-term evaluation plus what
// follows it before the next one: device copies, the Runge-Kutta state
// update, NSE::linear and the tau solve. Every algorithm step is a fixed
// number of ticks (SBDF: 1, SMRK2/CNRK2: 3, an initialisation step: 1 push
// + the init algorithm's ticks).
//
// An Integrator holds the state of one integration and produces its ticks; an
// Exec runs the ticks of any number of integrations on one grid as batched
// launches on its stream. A single DNS is an Exec with one Integrator; a batch
// is an Exec per grid class with many.
#pragma once

#include <unordered_map>
#include <unordered_set>

#include "cfg_host.hpp"

namespace cfg {

// Work of one integration for one tick, in execution order.
struct TickOps {
    struct Copy {
        double* dst;
        const double* src;
        size_t n;  // doubles
    };
    int ncopy = 0;
    Copy copy[4];
    bool nlon = false;    // a nonlinear-term evaluation (every time-step tick has one)
    int method = NL_ROT;  // its form
    NLMember nl{};
    bool comb = false;  // out = sum_j ccoef[j] cin[j] over one 3-component field
    int cn = 0;
    double ccoef[MAX_TERMS];
    const double* cin[MAX_TERMS];
    double* cout = nullptr;
    bool lin = false;
    LinearMember linm{};
    bool solve = false;
    TauMember tau{};

    void add_copy(double* dst, const double* src, size_t n) { copy[ncopy++] = Copy{dst, src, n}; }
};

class Exec;

// Velocity (3 components) and pressure, compact layout, pooled device memory.
struct FieldPair {
    PBuf u, q;
    FieldPair() = default;
    FieldPair(DevPool& pool, const Geom& g)
        : u(pool, 2L * 3 * g.Ny * g.M), q(pool, 2L * g.Ny * g.M) {}
};

// Current state: the integration's own buffers, or the front of a multistep
// history (which then holds the state, as after dnsalgo.cpp:207).
struct StateRef {
    double* u;
    double* q;
};

// One NSE object (nse.cpp): flow constants, nonlinear-form state, device
// scalars of the bulk-velocity solve, tau data per lambda_t.
class NSECtx {
   public:
    NSECtx(Exec& ex, std::shared_ptr<FlowModel> model, const cfg_flags& f);
    void reset_lambda(const std::vector<double>& lambda_t);
    void nonlinear(const double* u, double* f, TickOps& ops);
    void linear(const double* u, const double* q, double* out, TickOps& ops);
    void solve(double* u, double* q, std::initializer_list<std::pair<double, const double*>> terms, int s,
               TickOps& ops);
    double dPdx();
    double dPdz();
    FlowModel& model() { return *model_; }
    const cfg_flags& flags() const { return flags_; }

    // the tau sets this NSE holds (device memory accounting)
    void tau_sets(std::unordered_set<const TauStore*>& out) const {
        for (auto& t : tau_) out.insert(t.get());
    }
   private:
    void fetch_dP();
    Exec& ex_;
    std::shared_ptr<FlowModel> model_;
    cfg_flags flags_;
    int method_;
    PBuf dP_;
    bool dP_dirty_ = false;
    double dPdx_ = 0, dPdz_ = 0;
    std::vector<std::shared_ptr<TauStore>> tau_;
};

// A time-stepping algorithm of dnsalgo.cpp in tick form.
class Stepper {
   public:
    Stepper(NSECtx& nse, const cfg_flags& f) : nse_(nse), flags_(f), t_(f.t0) {}
    virtual ~Stepper() = default;
    virtual bool full() const = 0;
    virtual int ticks_per_step() const = 0;
    virtual int Ninitsteps() const = 0;
    virtual void tick(int sub, StateRef& cur, TickOps& ops) = 0;  // substep sub of one step
    virtual void push(StateRef& cur, TickOps& ops) {}              // init-phase history push
    virtual void reset_dt(double dt) = 0;
    virtual FieldPair* front() { return nullptr; }  // multistep history front, order > 1
    double time() const { return t_; }
    void reset_time(double t) { t_ = t; }

   protected:
    NSECtx& nse_;
    cfg_flags flags_;
    double t_;
};

std::unique_ptr<Stepper> make_stepper(Exec& ex, NSECtx& nse, const cfg_flags& f);

// One DNS integration (dns.cpp:22-163).
class Integrator {
   public:
    Integrator(Exec& ex, const cfg_flags& flags, const double* u_spec, const double* q_spec, const double* Ubase,
               const double* Wbase);
    void next(TickOps& ops);           // produce the next tick (host state advances)
    long steps() const { return steps_; }
    double time() const { return main_->time(); }
    StateRef state() const { return cur_; }
    void set_state(const double* u_spec, const double* q_spec);
    void get_state(double* u_spec, double* q_spec);
    void reset_dt(double dt);
    void reset_time(double t);
    void reset_gradp(double dPdx, double dPdz);
    void reset_bulkv(double Ubulk, double Wbulk);
    NLMember cfl_member() const;  // the current velocity, for Exec::cfl
    double cfl_scale() const;     // DNSAlgorithm::CFL's factor on CFLfactor (dnsalgo.cpp:63-67)
    double dPdx() { return main_nse_->dPdx(); }
    double Ubulk() const { return model_->UbulkAct(); }
    const FlowModel& model() const { return *model_; }
    const cfg_flags& flags() const { return flags_; }
    Exec& exec() { return ex_; }
    void tau_sets(std::unordered_set<const TauStore*>& out) const {
        if (main_nse_) main_nse_->tau_sets(out);
        if (init_nse_) init_nse_->tau_sets(out);
    }

   private:
    void own_state(TickOps* ops);  // make cur_ the integration's own buffers
    Exec& ex_;
    cfg_flags flags_;
    std::shared_ptr<FlowModel> model_;
    std::unique_ptr<NSECtx> main_nse_, init_nse_;
    std::unique_ptr<Stepper> main_, init_;
    FieldPair own_;
    StateRef cur_{};
    int init_sub_ = -1;  // -1: the next init-phase tick is the push
    int main_sub_ = 0;
    long steps_ = 0;
};

// One launch of a tick, with its tables already resident on the device.
struct LaunchRec {
    enum Kind { COPY, NONLINEAR, COMBINE, LINEAR, TAU } kind;
    int method, n, rotating, cn;
    NLTune tn;
    const void* tab;
    void* dst;
    const void* src;
    size_t bytes;
    double ccoef[MAX_TERMS];
    const double* cin[MAX_TERMS];
    double* cout;
};

// Runs ticks of integrations on one grid, batched, on one stream.
//
// With graphs on, a tick whose launch sequence (kernels, arguments and table
// addresses) has been seen before is captured into a hipGraph on its second
// occurrence and replayed from then on. Multistep histories rotate with a
// fixed period, so a class of members in steady state cycles through a few
// sequences. Graphs are dropped when the work buffers are reallocated or the
// launch-table arena is recycled, the two events that move an address a graph
// holds.
class Exec {
   public:
    explicit Exec(const cfg_grid& grid, bool dealias);
    ~Exec();
    Exec(const Exec&) = delete;
    Exec& operator=(const Exec&) = delete;
    void run(const std::vector<TickOps*>& ops);
    // CFL factors (flowfield.cpp:4035-4068) of the given fields, synchronously
    std::vector<double> cfl(const std::vector<NLMember>& fields);
    // for each field (n doubles), whether it holds a NaN or an infinity, synchronously
    std::vector<bool> nonfinite(const std::vector<const double*>& fields, long n);
    void sync() { HIPCHK(hipStreamSynchronize(st)); }
    void set_graphs(bool on);
    long graph_replays() const { return replays_; }
    long graph_captures() const { return captures_; }

    // A tick planned without being launched, for graphs that span several
    // Execs (Batch): prepare() makes the tables resident and fills the launch
    // records, append_key() adds them to a graph key, issue() launches them on
    // st (also inside a capture that st has joined). generation() changes when
    // an address a recorded launch holds may have moved.
    void prepare(const std::vector<TickOps*>& ops);
    void append_key(std::string& key) const;
    unsigned long generation() const { return buf_gen_ * 1000003UL + cache.generation(); }
    void issue();  // launches recs_ on st

    // Device memory: the work buffers hold cap() members; a tick of n members
    // grows them to grown_cap(cap(), n). held_bytes() is what this Exec holds
    // for its integrations (pool blocks in use or free, work buffers; not its
    // launch-table arena); trim() hands the pool's free blocks and the work
    // buffers back to HIP.
    int cap() const { return cap_; }
    static int grown_cap(int cap, int n) { return n <= cap ? cap : std::max(n, 2 * cap); }
    static size_t work_bytes(const Geom& g, int cap);
    size_t held_bytes() const;
    void trim();

    std::shared_ptr<Geometry> geo;
    hipStream_t st = nullptr;
    DevPool pool;  // before everything that returns memory to it
    TableCache cache;
    KTimer kt;
    long nfield = 0;  // doubles in one 3-component compact field
    long nscal = 0;   // doubles in one 1-component compact field

   private:
    void reserve(int n);
    void plan(const std::vector<TickOps*>& ops);  // fills recs_, every table of one arena generation
    void plan_once(const std::vector<TickOps*>& ops);
    void drop_graphs();
    int cap_ = 0;
    DBuf yp_, xp_, zo_, fo_, cflout_, nfout_;
    std::vector<NLMember> nlt_;
    std::vector<LinearMember> lt_;
    std::vector<TauMember> tt_;
    std::vector<LaunchRec> recs_;
    std::map<std::pair<int, int>, NLTune> tune_;  // (nonlinear form, member-count bucket) -> tiling
    bool graphs_on_ = false;
    unsigned long buf_gen_ = 0, graph_gen_ = ~0UL;  // buffer generation; generation the graphs belong to
    std::unordered_map<std::string, hipGraphExec_t> graphs_;
    std::unordered_set<std::string> seen_;
    long replays_ = 0, captures_ = 0;
};

}  // namespace cfg
