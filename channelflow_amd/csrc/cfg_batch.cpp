// Batched and served DNS integrations (include/cfgpu.h, cfg_batch_* and
// cfg_server_*). channelflow runs its many independent integrations one
// after another (edge-tracking bisection members, multi-shooting segments,
// DSI Jacobian probes: cfdsi.cpp:678-775, multiShootingDSI.cpp:70-86,
// edgetracking.cpp:472-530); here they advance together.
//
// Integrations on the same grid (and dealiasing) form a class. A class owns
// one Exec (stream, work buffers, launch-table cache, memory pool) and runs
// one tick of all its members per launch sequence; classes run on separate
// streams and overlap on the device. Members never exchange data, and every
// kernel treats each member on its own, so a member's result is the same alone
// and in any mix.
#include <chrono>
#include <deque>

#include "cfg_step.hpp"

namespace cfg {

struct BatchClass {
    std::unique_ptr<Exec> ex;
    std::vector<int> members;  // ids, in insertion order
    std::deque<TickOps> ops;   // this tick's operations of the members below their target
    std::vector<TickOps*> ptrs;
    hipEvent_t done = nullptr;  // marks the class stream's work for a whole-batch launch
    ~BatchClass() {
        if (done) (void)hipEventDestroy(done);
    }
};

struct BatchMember {
    int id;
    BatchClass* cls;
    std::unique_ptr<Integrator> it;
    long target = 0;  // steps to reach in the current advance / serving horizon
    bool finite = true;  // state finite after its last advance
};

// hipGraph replay of repeating ticks: none (eager launches), one graph per
// grid class and tick, each class on its own stream (the default of a Batch),
// or one graph per tick for the whole batch, its classes captured as parallel
// branches forked from the batch's stream (one launch per tick, but more host
// time per tick on ROCm, with the same device time). A server's membership
// changes every few ticks, and its keys with it, so the server launches
// eagerly.
enum GraphMode { GRAPHS_OFF = 0, GRAPHS_PER_CLASS = 1, GRAPHS_WHOLE = 2 };

class Batch {
   public:
    Batch() {
        HIPCHK(hipStreamCreate(&bst_));
        HIPCHK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&after_, hipEventDisableTiming));
    }
    ~Batch() {
        for (auto& kv : classes_) kv.second->ex->sync();
        (void)hipStreamSynchronize(bst_);
        drop_whole();
        members_.clear();  // integrators before their Exec
        classes_.clear();
        (void)hipEventDestroy(fork_);
        (void)hipEventDestroy(after_);
        (void)hipStreamDestroy(bst_);
    }

    int add(const cfg_grid& grid, const cfg_flags& flags, const double* u, const double* q) {
        check_flags(flags);
        check_grid(grid);
        const GridKey k = grid_key(grid, flags.dealiasing == CFG_DealiasXZ);
        auto cit = classes_.find(k);
        if (cit == classes_.end()) {
            auto c = std::make_unique<BatchClass>();
            c->ex = std::make_unique<Exec>(grid, flags.dealiasing == CFG_DealiasXZ);
            c->ex->set_graphs(mode_ == GRAPHS_PER_CLASS);
            HIPCHK(hipEventCreateWithFlags(&c->done, hipEventDisableTiming));
            cit = classes_.emplace(k, std::move(c)).first;
        }
        BatchClass* cls = cit->second.get();
        auto m = std::make_unique<BatchMember>();
        m->id = next_id_++;
        m->cls = cls;
        m->it = std::make_unique<Integrator>(*cls->ex, flags, u, q, nullptr, nullptr);
        cls->members.push_back(m->id);
        const int id = m->id;
        members_.emplace(id, std::move(m));
        return id;
    }

    void remove(int id) {
        BatchMember& m = get(id);
        auto& v = m.cls->members;
        v.erase(std::find(v.begin(), v.end(), id));
        members_.erase(id);  // its buffers return to the class pool (stream-ordered reuse)
    }

    BatchMember& get(int id) {
        auto it = members_.find(id);
        if (it == members_.end()) throw Error(CFG_ERR_ARG, "no such batch member");
        return *it->second;
    }
    int size() const { return (int)members_.size(); }

    // One tick of every member below its target. Returns false when none was.
    bool tick() {
        const auto t0 = std::chrono::steady_clock::now();
        active_.clear();
        bool timing = false;
        for (auto& kv : classes_) {
            BatchClass& c = *kv.second;
            c.ops.clear();
            c.ptrs.clear();
            for (int id : c.members) {
                BatchMember& m = *members_.at(id);
                if (m.it->steps() < m.target) {
                    c.ops.emplace_back();
                    m.it->next(c.ops.back());
                }
            }
            if (c.ops.empty()) continue;
            for (auto& o : c.ops) c.ptrs.push_back(&o);
            active_.push_back(&c);
            timing = timing || c.ex->kt.on;
        }
        if (active_.empty()) return false;
        ++ticks_;
        if (mode_ != GRAPHS_WHOLE || timing) {  // eager, or each class's own graphs
            for (BatchClass* c : active_) c->ex->run(c->ptrs);
        } else {
            whole_tick();
        }
        host_us_ += std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }
    void advance(int n) {
        for (auto& kv : members_) kv.second->target = kv.second->it->steps() + n;
        while (tick()) {
        }
        sync();
        if (n > 0) {
            std::vector<int> ids;
            for (auto& kv : members_) ids.push_back(kv.first);
            check_finite(ids);
        }
    }

    // the isfinite(L2Norm(u)) test of cfdsi.cpp:769-774 on the listed members'
    // states, one launch per class; throws naming the members that fail (their
    // states stay as they are)
    void check_finite(const std::vector<int>& ids) {
        std::map<BatchClass*, std::vector<BatchMember*>> by;
        for (int id : ids) {
            BatchMember& m = get(id);
            by[m.cls].push_back(&m);
        }
        std::string bad;
        for (auto& kv : by) {
            std::vector<const double*> f;
            for (BatchMember* m : kv.second) f.push_back(m->it->state().u);
            const std::vector<bool> nf = kv.first->ex->nonfinite(f, kv.first->ex->nfield);
            for (size_t i = 0; i < nf.size(); ++i) {
                kv.second[i]->finite = !nf[i];
                if (nf[i]) bad += (bad.empty() ? "" : ", ") + std::to_string(kv.second[i]->id);
            }
        }
        if (!bad.empty()) throw Error(CFG_ERR_NONFINITE, "velocity not finite after advance: member(s) " + bad);
    }

    // DNS::advance(fields, nsteps[k]) of the listed members, together; the
    // other members stay where they are
    void advance_members(int count, const int* ids, const int* nsteps) {
        for (int k = 0; k < count; ++k) {
            (void)get(ids[k]);
            if (nsteps[k] < 0) throw Error(CFG_ERR_ARG, "negative step count");
        }
        for (auto& kv : members_) kv.second->target = kv.second->it->steps();
        for (int k = 0; k < count; ++k) get(ids[k]).target += nsteps[k];
        while (tick()) {
        }
        sync();
        std::vector<int> moved;
        for (int k = 0; k < count; ++k)
            if (nsteps[k] > 0) moved.push_back(ids[k]);
        check_finite(moved);
    }

    void sync() {
        HIPCHK(hipStreamSynchronize(bst_));
        for (auto& kv : classes_) kv.second->ex->sync();
    }

    void profile(bool on) {
        for (auto& kv : classes_) {
            kv.second->ex->kt.reset();
            kv.second->ex->kt.on = on;
        }
    }
    void kernel_times(double* ms, long long* cnt) {
        for (int i = 0; i < K_COUNT; ++i) {
            if (ms) ms[i] = 0.0;
            if (cnt) cnt[i] = 0;
        }
        for (auto& kv : classes_) {
            KTimer& kt = kv.second->ex->kt;
            kt.flush();
            for (int i = 0; i < K_COUNT; ++i) {
                if (ms) ms[i] += kt.total_ms[i];
                if (cnt) cnt[i] += kt.count[i];
            }
        }
    }
    int nclasses() const { return (int)classes_.size(); }

    BatchClass* find_class(const cfg_grid& grid, bool dealias) {
        auto it = classes_.find(grid_key(grid, dealias));
        return it == classes_.end() ? nullptr : it->second.get();
    }
    // Device memory held for the integrations: every class's pool blocks (in
    // use or free) and work buffers, and the tau sets of the members (each
    // once). The classes' fixed tables are not counted.
    size_t held_bytes() const {
        size_t b = 0;
        for (auto& kv : classes_) b += kv.second->ex->held_bytes();
        std::unordered_set<const TauStore*> tau;
        for (auto& kv : members_) kv.second->it->tau_sets(tau);
        for (const TauStore* t : tau) b += t->bytes;
        return b;
    }
    // a class left without members hands its free blocks and work buffers
    // back to HIP
    void trim_if_empty(BatchClass* c) {
        if (c && c->members.empty()) c->ex->trim();
    }

    // hipGraph capture and replay of repeating ticks (GraphMode; one graph
    // per class by default: ROCm's launch of a graph whose classes are
    // parallel branches cost 223 us of host time per tick for the three C3
    // classes, and the class-stream ordering events 130 us more, against
    // 3 x 15 us for three single-stream graphs, with the same GPU time)
    void set_graphs(int mode) {
        if (mode < GRAPHS_OFF || mode > GRAPHS_WHOLE) throw Error(CFG_ERR_ARG, "graph mode must be 0, 1 or 2");
        sync();
        drop_whole();
        mode_ = (GraphMode)mode;
        for (auto& kv : classes_) kv.second->ex->set_graphs(mode_ == GRAPHS_PER_CLASS);
    }
    void graph_stats(long long* captures, long long* replays) const {
        long long c = captures_, r = replays_;
        for (auto& kv : classes_) {
            c += kv.second->ex->graph_captures();
            r += kv.second->ex->graph_replays();
        }
        if (captures) *captures = c;
        if (replays) *replays = r;
    }
    // ticks so far, graph launches among them (replays and launches right
    // after a capture), and host microseconds spent planning and launching
    void tick_stats(long long* ticks, long long* graph_launches, double* host_us) const {
        long long gl = replays_ + captures_;
        for (auto& kv : classes_) gl += kv.second->ex->graph_replays() + kv.second->ex->graph_captures();
        if (ticks) *ticks = ticks_;
        if (graph_launches) *graph_launches = gl;
        if (host_us) *host_us = host_us_;
    }

   private:
    // One tick of the active classes as one graph on bst_. The key is every
    // active class's identity, generation and launch records; a key seen for
    // the second time is captured, with each class's launches on its own
    // stream forked from bst_ and joined back, so the classes stay parallel
    // branches of the graph. Class streams and bst_ are ordered with events
    // around every whole-batch launch, so work a class stream takes outside
    // these graphs (uploads, copies, checks, frees) stays in order.
    void whole_tick() {
        std::string key;
        for (BatchClass* c : active_) {
            c->ex->prepare(c->ptrs);
            const unsigned long gen = c->ex->generation();
            key.append(reinterpret_cast<const char*>(&c), sizeof c);
            key.append(reinterpret_cast<const char*>(&gen), sizeof gen);
            c->ex->append_key(key);
        }
        auto it = whole_.find(key);
        if (it != whole_.end()) {
            launch_whole(it->second);
            ++replays_;
            return;
        }
        if (!seen_.count(key)) {  // first occurrence: eager, each class on its stream
            if (seen_.size() + whole_.size() >= kMaxWholeKeys) {  // membership churn: start over
                sync();
                drop_whole();
            }
            seen_.insert(key);
            for (BatchClass* c : active_) c->ex->issue();
            return;
        }
        join_classes();  // earlier class-stream work before the capture's fork
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(bst_, hipStreamCaptureModeThreadLocal));
        try {
            HIPCHK(hipEventRecord(fork_, bst_));
            for (BatchClass* c : active_) {
                HIPCHK(hipStreamWaitEvent(c->ex->st, fork_, 0));
                c->ex->issue();
                HIPCHK(hipEventRecord(c->done, c->ex->st));
                HIPCHK(hipStreamWaitEvent(bst_, c->done, 0));
            }
        } catch (...) {
            (void)hipStreamEndCapture(bst_, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        HIPCHK(hipStreamEndCapture(bst_, &gr));
        hipGraphExec_t ge = nullptr;
        const hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
        seen_.erase(key);
        whole_.emplace(std::move(key), ge);
        ++captures_;
        launch_whole(ge);
    }
    // bst_ waits for every class stream's work so far
    void join_classes() {
        for (auto& kv : classes_) {
            HIPCHK(hipEventRecord(kv.second->done, kv.second->ex->st));
            HIPCHK(hipStreamWaitEvent(bst_, kv.second->done, 0));
        }
    }
    void launch_whole(hipGraphExec_t ge) {
        join_classes();
        HIPCHK(hipGraphLaunch(ge, bst_));
        HIPCHK(hipEventRecord(after_, bst_));
        for (auto& kv : classes_) HIPCHK(hipStreamWaitEvent(kv.second->ex->st, after_, 0));
    }
    void drop_whole() {
        for (auto& kv : whole_) (void)hipGraphExecDestroy(kv.second);
        whole_.clear();
        seen_.clear();
    }

    static constexpr size_t kMaxWholeKeys = 256;
    GraphMode mode_ = GRAPHS_PER_CLASS;
    hipStream_t bst_ = nullptr;
    hipEvent_t fork_ = nullptr, after_ = nullptr;
    std::unordered_map<std::string, hipGraphExec_t> whole_;
    std::unordered_set<std::string> seen_;
    std::vector<BatchClass*> active_;
    long long captures_ = 0, replays_ = 0, ticks_ = 0;
    double host_us_ = 0.0;
    std::map<GridKey, std::unique_ptr<BatchClass>> classes_;
    std::map<int, std::unique_ptr<BatchMember>> members_;
    int next_id_ = 0;
};

// The device memory of an integration: the pool blocks (doubles each) its
// Integrator and steppers make (cfg_step.cpp: its state u and q; a multistep
// scheme's history of u and q and of f; the three field buffers of a
// CNAB-style or Runge-Kutta scheme, as the main one or to start a multistep
// one; the mean pressure gradients of the main and start-up NSE) and its tau
// sets (one lambda each: 1 for a multistep or CNAB2 scheme, 3 for SMRK2 and
// CNRK2, plus those of the start-up scheme).
struct MemberPlan {
    std::vector<size_t> blocks;
    int tau_sets = 0;
};
inline MemberPlan member_plan(long Ny, long M, const cfg_flags& f) {
    const size_t nf = 2 * 3 * (size_t)Ny * M, ns = 2 * (size_t)Ny * M;
    MemberPlan p;
    p.blocks = {nf, ns, 2, 2};
    auto three = [&] {
        for (int i = 0; i < 3; ++i) p.blocks.push_back(nf);
    };
    int order = 0;
    switch (f.timestepping) {
        case CFG_SBDF1: case CFG_CNFE1: order = 1; break;
        case CFG_SBDF2: order = 2; break;
        case CFG_SBDF3: order = 3; break;
        case CFG_SBDF4: order = 4; break;
        case CFG_CNAB2: three(); p.tau_sets = 1; break;
        default: three(); p.tau_sets = 3; break;  // SMRK2, CNRK2
    }
    if (order > 0) {
        p.tau_sets = 1;
        for (int j = 0; j < order; ++j) {
            if (order > 1) {
                p.blocks.push_back(nf);
                p.blocks.push_back(ns);
            }
            p.blocks.push_back(nf);
        }
        if (order > 1) {  // the start-up scheme
            three();
            p.tau_sets += f.initstepping == CFG_CNAB2 ? 1 : 3;
        }
    }
    return p;
}

// Plans admissions against the device memory a batch holds for its
// integrations: a request needs the pool blocks its class's free lists cannot
// supply, its tau sets (counted in full, though members with the same grid,
// nu and dt share them) and the growth of its class's work buffers
// (Exec::grown_cap). A class's fixed tables (launch-table arena, geometry)
// are outside the budget.
class AdmissionPlan {
   public:
    // b null: plan against an empty batch (the request alone)
    explicit AdmissionPlan(Batch* b) : batch_(b), held_(b ? b->held_bytes() : 0) {}
    double held() const { return (double)held_; }
    double committed() const { return (double)(held_ + pending_); }
    size_t need(const cfg_grid& gr, const cfg_flags& f, bool commit) {
        const bool de = f.dealiasing == CFG_DealiasXZ;
        const long kxd = de ? gr.Nx / 3 - 1 : gr.Nx / 2 - 1, kzd = de ? gr.Nz / 3 - 1 : gr.Nz / 2 - 1;
        Geom g{};
        g.Nx = gr.Nx;
        g.Ny = gr.Ny;
        g.Nz = gr.Nz;
        g.nkx = (int)(2 * kxd + 1);
        g.nkz = (int)(kzd + 1);
        g.M = g.nkx * g.nkz;
        const GridKey k = grid_key(gr, de);
        BatchClass* cls = batch_ ? batch_->find_class(gr, de) : nullptr;
        size_t bytes = 0;
        const MemberPlan mp = member_plan(g.Ny, g.M, f);
        std::map<size_t, int> use;
        for (size_t n : mp.blocks) {
            const int avail = cls ? (int)cls->ex->pool.free_count(n) - claimed_[{k, n}] - use[n] : 0;
            if (avail > 0) ++use[n];
            else bytes += device_bytes_taken(n * sizeof(double));
        }
        bytes += (size_t)mp.tau_sets * tau_set_bytes(g);
        bytes += kMemberTables;
        auto cit = cap_.find(k);
        const int cap0 = cit != cap_.end() ? cit->second : (cls ? cls->ex->cap() : 0);
        const int n0 = members_.count(k) ? members_[k] : (cls ? (int)cls->members.size() : 0);
        const int cap1 = Exec::grown_cap(cap0, n0 + 1);
        bytes += Exec::work_bytes(g, cap1) - Exec::work_bytes(g, cap0);
        if (commit) {
            for (auto& kv : use) claimed_[{k, kv.first}] += kv.second;
            cap_[k] = cap1;
            members_[k] = n0 + 1;
            pending_ += bytes;
        }
        return bytes;
    }

   private:
    // the small device tables of an integration's model (base-flow values and
    // constants of its solves), each its own allocation
    static constexpr size_t kMemberTables = size_t(64) << 10;
    Batch* batch_;
    size_t held_, pending_ = 0;
    std::map<std::pair<GridKey, size_t>, int> claimed_;
    std::map<GridKey, int> cap_, members_;
};

// Continuous batching of integration requests: queued requests are admitted
// while the device-memory budget allows, every admitted request advances one
// tick per server step alongside the others, and a request that reaches its
// horizon leaves at once (its slot and memory go to the next in the queue).
// Results come back through pinned host memory without stalling the others.
class Server {
   public:
    // Eager launches: the membership changes every few ticks as requests come
    // and go, so a captured tick is rarely replayed often enough to pay for its
    // capture (96 requests, Poisson 8/s: p50 0.42 s eager against 0.51 s with
    // per-class graphs, 4.97e9 against 4.95e9 grid-pt*steps/s; all queued at
    // t = 0: 6.43e9 against 6.38e9).
    explicit Server(double budget) : budget_(budget) { batch_.set_graphs(GRAPHS_OFF); }
    ~Server() {
        batch_.sync();
        for (auto& kv : done_)
            if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
        for (auto& kv : done_)
            if (kv.second.host) (void)hipHostFree(kv.second.host);
    }

    long long submit(const cfg_grid& g, const cfg_flags& f, const double* u, const double* q, int nsteps) {
        check_flags(f);
        check_grid(g);
        if (nsteps < 0) throw Error(CFG_ERR_ARG, "negative horizon");
        const double bytes = (double)AdmissionPlan(nullptr).need(g, f, false);
        if (bytes > budget_) throw Error(CFG_ERR_ARG, "request does not fit the device-memory budget");
        Req r;
        r.id = next_id_++;
        r.grid = g;
        r.flags = f;
        r.nsteps = nsteps;
        r.bytes = bytes;
        const size_t nu = (size_t)2 * 3 * g.Ny * g.Nx * (g.Nz / 2 + 1), nq = nu / 3;
        r.u.assign(u, u + nu);
        if (q) r.q.assign(q, q + nq);
        queue_.push_back(std::move(r));
        return queue_.back().id;
    }

    // One server step: admit, one tick of every running request, retire the
    // ones at their horizon, report results that have reached the host.
    void step(long long* done_ids, int max_done, int* ndone, int* nrun, int* nqueued) {
        admit();
        batch_.tick();
        for (auto it = running_.begin(); it != running_.end();) {
            BatchMember& m = batch_.get(it->second.member);
            if (m.it->steps() >= m.target) {
                retire(it->first, it->second);
                it = running_.erase(it);
            } else {
                ++it;
            }
        }
        int n = 0;
        for (auto& kv : done_) {
            Done& d = kv.second;
            if (d.reported || !d.ev) continue;
            if (hipEventQuery(d.ev) != hipSuccess) continue;
            d.ready = true;
            if (n < max_done) {
                if (done_ids) done_ids[n] = kv.first;
                d.reported = true;
                ++n;
            }
        }
        if (ndone) *ndone = n;
        if (nrun) *nrun = (int)running_.size();
        if (nqueued) *nqueued = (int)queue_.size();
    }

    void result(long long id, double* u_spec, double* q_spec, double* t) {
        auto it = done_.find(id);
        if (it == done_.end()) throw Error(CFG_ERR_ARG, "no such completed request");
        Done& d = it->second;
        HIPCHK(hipEventSynchronize(d.ev));
        std::vector<double> hu(d.host, d.host + d.nu), hq(d.host + d.nu, d.host + d.nu + d.nq);
        if (u_spec) unpack(d.g, 3, hu, u_spec);
        if (q_spec) unpack(d.g, 1, hq, q_spec);
        if (t) *t = d.t;
        (void)hipEventDestroy(d.ev);
        (void)hipHostFree(d.host);
        done_.erase(it);
        // cfdsi.cpp:769-774: a non-finite f^T(u) is an error (the result is still handed out)
        for (double v : hu)
            if (!std::isfinite(v))
                throw Error(CFG_ERR_NONFINITE, "request " + std::to_string(id) + ": velocity not finite at its horizon");
    }

    // device memory the server holds (its batch's pools, work buffers, table
    // arenas and the tau data)
    double used() const { return (double)batch_.held_bytes(); }

   private:
    struct Req {
        long long id;
        cfg_grid grid;
        cfg_flags flags;
        int nsteps;
        double bytes;
        std::vector<double> u, q;
    };
    struct Run {
        int member;
        double bytes;
    };
    struct Done {
        Geom g;
        double* host = nullptr;
        size_t nu = 0, nq = 0;
        hipEvent_t ev = nullptr;
        double t = 0;
        bool ready = false, reported = false;
    };

    // Admission in arrival order, with backfilling: when the request at the
    // head of the queue does not fit the free budget, later requests that fit
    // are admitted past it, until the head has waited kMaxHeadWait server
    // steps; from then on nothing passes it, so running requests drain until
    // it fits (a large request is delayed by small ones at most that long).
    // What fits is planned against the memory the batch holds now plus what
    // the requests admitted in this step will take (AdmissionPlan).
    void admit() {
        if (queue_.empty()) return;
        AdmissionPlan plan(&batch_);
        auto fits = [&](const Req& r) {
            return plan.committed() + (double)plan.need(r.grid, r.flags, false) <= budget_;
        };
        if (!fits(queue_.front())) ++head_wait_;
        for (auto it = queue_.begin(); it != queue_.end();) {
            const bool head = it == queue_.begin();
            if (fits(*it)) {
                Req r = std::move(*it);
                it = queue_.erase(it);
                plan.need(r.grid, r.flags, true);
                const int mid = batch_.add(r.grid, r.flags, r.u.data(), r.q.empty() ? nullptr : r.q.data());
                batch_.get(mid).target = r.nsteps;
                running_.emplace(r.id, Run{mid, r.bytes});
                if (head) head_wait_ = 0;
                continue;
            }
            if (head_wait_ > kMaxHeadWait) break;  // the head has waited long enough: no more passing
            ++it;
        }
    }
    static constexpr long kMaxHeadWait = 256;
    long head_wait_ = 0;  // server steps the current head of the queue has been blocked

    void retire(long long id, const Run& run) {
        BatchMember& m = batch_.get(run.member);
        Exec& ex = *m.cls->ex;
        Done d;
        d.g = ex.geo->g;
        d.nu = ex.nfield;
        d.nq = ex.nscal;
        d.t = m.it->time();
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&d.host), (d.nu + d.nq) * sizeof(double)));
        const StateRef s = m.it->state();
        HIPCHK(hipMemcpyAsync(d.host, s.u, d.nu * sizeof(double), hipMemcpyDeviceToHost, ex.st));
        HIPCHK(hipMemcpyAsync(d.host + d.nu, s.q, d.nq * sizeof(double), hipMemcpyDeviceToHost, ex.st));
        HIPCHK(hipEventCreateWithFlags(&d.ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(d.ev, ex.st));
        BatchClass* cls = m.cls;
        batch_.remove(run.member);  // buffers go back to the class pool, reused only by later work
        batch_.trim_if_empty(cls);  // after the copy above (trim synchronizes the class stream)
        // tau sets the request alone used go back to HIP now (hipFree waits for
        // the kernels that used them), so the budget's count stays the device's
        deferred_free(nullptr, 0, true);
        done_.emplace(id, d);
    }

    Batch batch_;
    double budget_;
    std::deque<Req> queue_;
    std::map<long long, Run> running_;
    std::map<long long, Done> done_;
    long long next_id_ = 0;
};

}  // namespace cfg

using namespace cfg;

struct cfg_batch {
    Batch b;
};
struct cfg_server {
    explicit cfg_server(double budget) : s(budget) {}
    Server s;
};

extern "C" {

int cfg_batch_create(cfg_batch** out) {
    return guarded([&] {
        if (!out) throw Error(CFG_ERR_ARG, "null argument");
        *out = new cfg_batch();
    });
}

int cfg_batch_destroy(cfg_batch* b) {
    return guarded([&] { delete b; });
}

int cfg_batch_add(cfg_batch* b, const cfg_grid* grid, const cfg_flags* flags, const double* u_spec,
                  const double* q_spec, int* member) {
    return guarded([&] {
        if (!b || !grid || !flags || !u_spec || !member) throw Error(CFG_ERR_ARG, "null argument");
        *member = b->b.add(*grid, *flags, u_spec, q_spec);
    });
}

int cfg_batch_remove(cfg_batch* b, int member) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.remove(member);
    });
}

int cfg_batch_advance(cfg_batch* b, int nsteps) {
    return guarded([&] {
        if (!b || nsteps < 0) throw Error(CFG_ERR_ARG, "bad argument");
        b->b.advance(nsteps);
    });
}

int cfg_batch_size(const cfg_batch* b, int* n, int* nclasses) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        if (n) *n = b->b.size();
        if (nclasses) *nclasses = b->b.nclasses();
    });
}

int cfg_batch_get_state(cfg_batch* b, int member, double* u_spec, double* q_spec) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.get(member).it->get_state(u_spec, q_spec);
    });
}

int cfg_batch_set_state(cfg_batch* b, int member, const double* u_spec, const double* q_spec) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.get(member).it->set_state(u_spec, q_spec);
    });
}

int cfg_batch_time(cfg_batch* b, int member, double* t) {
    return guarded([&] {
        if (!b || !t) throw Error(CFG_ERR_ARG, "null argument");
        *t = b->b.get(member).it->time();
    });
}

int cfg_batch_finite(cfg_batch* b, int member, int* finite) {
    return guarded([&] {
        if (!b || !finite) throw Error(CFG_ERR_ARG, "null argument");
        *finite = b->b.get(member).finite ? 1 : 0;
    });
}

int cfg_batch_dpdx(cfg_batch* b, int member, double* v) {
    return guarded([&] {
        if (!b || !v) throw Error(CFG_ERR_ARG, "null argument");
        *v = b->b.get(member).it->dPdx();
    });
}

int cfg_batch_advance_members(cfg_batch* b, int count, const int* members, const int* nsteps) {
    return guarded([&] {
        if (!b || count < 0 || (count > 0 && (!members || !nsteps))) throw Error(CFG_ERR_ARG, "bad argument");
        b->b.advance_members(count, members, nsteps);
    });
}

int cfg_batch_reset_dt(cfg_batch* b, int member, double dt) {  // DNS::reset_dt, dns.cpp:180-185
    return guarded([&] {
        if (!b || !(dt > 0)) throw Error(CFG_ERR_ARG, "bad argument");
        BatchMember& m = b->b.get(member);
        m.it->reset_dt(dt);
        m.cls->ex->sync();
    });
}

int cfg_batch_profile(cfg_batch* b, int enable) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.profile(enable != 0);
    });
}

int cfg_batch_kernel_times(cfg_batch* b, double* total_ms, long long* counts) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.kernel_times(total_ms, counts);
    });
}

int cfg_batch_cfl(cfg_batch* b, int member, double* cfl) {
    return guarded([&] {
        if (!b || !cfl) throw Error(CFG_ERR_ARG, "null argument");
        BatchMember& m = b->b.get(member);
        *cfl = m.cls->ex->cfl({m.it->cfl_member()})[0] * m.it->cfl_scale();
    });
}

int cfg_batch_set_graphs(cfg_batch* b, int enable) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.set_graphs(enable);
    });
}

int cfg_batch_graph_stats(const cfg_batch* b, long long* captures, long long* replays) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.graph_stats(captures, replays);
    });
}

int cfg_batch_tick_stats(const cfg_batch* b, long long* ticks, long long* graph_launches, double* host_us) {
    return guarded([&] {
        if (!b) throw Error(CFG_ERR_ARG, "null batch");
        b->b.tick_stats(ticks, graph_launches, host_us);
    });
}

int cfg_server_create(double budget_bytes, cfg_server** out) {
    return guarded([&] {
        if (!out || !(budget_bytes > 0)) throw Error(CFG_ERR_ARG, "bad argument");
        *out = new cfg_server(budget_bytes);
    });
}

int cfg_server_destroy(cfg_server* s) {
    return guarded([&] { delete s; });
}

int cfg_server_submit(cfg_server* s, const cfg_grid* grid, const cfg_flags* flags, const double* u_spec,
                      const double* q_spec, int nsteps, long long* request) {
    return guarded([&] {
        if (!s || !grid || !flags || !u_spec || !request) throw Error(CFG_ERR_ARG, "null argument");
        *request = s->s.submit(*grid, *flags, u_spec, q_spec, nsteps);
    });
}

int cfg_server_step(cfg_server* s, long long* done, int max_done, int* ndone, int* nrunning, int* nqueued) {
    return guarded([&] {
        if (!s || max_done < 0) throw Error(CFG_ERR_ARG, "bad argument");
        s->s.step(done, max_done, ndone, nrunning, nqueued);
    });
}

int cfg_server_result(cfg_server* s, long long request, double* u_spec, double* q_spec, double* t) {
    return guarded([&] {
        if (!s) throw Error(CFG_ERR_ARG, "null server");
        s->s.result(request, u_spec, q_spec, t);
    });
}

int cfg_server_memory(const cfg_server* s, double* used_bytes) {
    return guarded([&] {
        if (!s || !used_bytes) throw Error(CFG_ERR_ARG, "null argument");
        *used_bytes = s->s.used();
    });
}

}  // extern "C"
