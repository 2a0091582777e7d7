// C ABI of the single-field DNS (include/cfgpu.h): channelflow's DNS object
// (dns.h:30-97) over the tick engine of cfg_step.cpp. One cfg_dns owns one HIP
// stream and all of its device buffers; the velocity, pressure and multistep
// history stay resident between calls.
#include "cfg_step.hpp"

using namespace cfg;

struct cfg_dns {
    std::unique_ptr<Exec> ex;         // destroyed after the integrator
    std::unique_ptr<Integrator> it;
    std::vector<TickOps*> one;
    TickOps ops;

    void advance(int N) {  // dns.cpp:135-163
        const long target = it->steps() + N;
        while (it->steps() < target) {
            it->next(ops);
            ex->run(one);
        }
    }
};

extern "C" {

const char* cfg_last_error(void) { return last_error().c_str(); }

int cfg_device_count(int* n) {
    return guarded([&] {
        if (!n) throw Error(CFG_ERR_ARG, "null argument");
        HIPCHK(hipGetDeviceCount(n));
    });
}

int cfg_dns_create(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, const double* q_spec,
                   const double* Ubase, const double* Wbase, cfg_dns** out) {
    return guarded([&] {
        if (!grid || !flags || !u_spec || !out) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        auto d = std::make_unique<cfg_dns>();
        d->ex = std::make_unique<Exec>(*grid, flags->dealiasing == CFG_DealiasXZ);
        d->it = std::make_unique<Integrator>(*d->ex, *flags, u_spec, q_spec, Ubase, Wbase);
        d->one = {&d->ops};
        *out = d.release();
    });
}

int cfg_dns_destroy(cfg_dns* dns) {
    return guarded([&] {
        if (dns) {
            dns->ex->sync();
            delete dns;
        }
    });
}

int cfg_dns_advance(cfg_dns* dns, int nsteps) {
    return guarded([&] {
        if (!dns || nsteps < 0) throw Error(CFG_ERR_ARG, "bad argument");
        dns->advance(nsteps);
        dns->ex->sync();
        // cfdsi.cpp:769-774 stops on a non-finite f^T(u); the state is kept for the caller
        if (nsteps > 0 && dns->ex->nonfinite({dns->it->state().u}, dns->ex->nfield)[0])
            throw Error(CFG_ERR_NONFINITE, "velocity not finite after advance (t = " +
                                               std::to_string(dns->it->time()) + ")");
    });
}

int cfg_dns_set_state(cfg_dns* dns, const double* u_spec, const double* q_spec) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->it->set_state(u_spec, q_spec);
    });
}

int cfg_dns_get_state(cfg_dns* dns, double* u_spec, double* q_spec) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->it->get_state(u_spec, q_spec);
    });
}

int cfg_dns_reset_dt(cfg_dns* dns, double dt) {  // dns.cpp:180-185
    return guarded([&] {
        if (!dns || !(dt > 0)) throw Error(CFG_ERR_ARG, "bad argument");
        dns->it->reset_dt(dt);
        dns->ex->sync();
    });
}

int cfg_dns_reset_time(cfg_dns* dns, double t) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->it->reset_time(t);
    });
}

int cfg_dns_reset_gradp(cfg_dns* dns, double dPdx, double dPdz) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->it->reset_gradp(dPdx, dPdz);
    });
}

int cfg_dns_reset_bulkv(cfg_dns* dns, double Ubulk, double Wbulk) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->it->reset_bulkv(Ubulk, Wbulk);
    });
}

int cfg_dns_time(const cfg_dns* dns, double* t) {
    return guarded([&] {
        if (!dns || !t) throw Error(CFG_ERR_ARG, "null argument");
        *t = dns->it->time();
    });
}

int cfg_dns_dpdx(cfg_dns* dns, double* v) {
    return guarded([&] {
        if (!dns || !v) throw Error(CFG_ERR_ARG, "null argument");
        *v = dns->it->dPdx();
    });
}

int cfg_dns_ubulk(const cfg_dns* dns, double* v) {
    return guarded([&] {
        if (!dns || !v) throw Error(CFG_ERR_ARG, "null argument");
        *v = dns->it->Ubulk();
    });
}

int cfg_dns_cfl(cfg_dns* dns, double* cfl) {  // DNS::CFL, dns.cpp:282-289
    return guarded([&] {
        if (!dns || !cfl) throw Error(CFG_ERR_ARG, "null argument");
        *cfl = dns->it->exec().cfl({dns->it->cfl_member()})[0] * dns->it->cfl_scale();
    });
}

int cfg_dns_base(const cfg_dns* dns, double* U, double* W) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        const auto& Ub = dns->it->model().Ubase();
        const auto& Wb = dns->it->model().Wbase();
        if (U) std::copy(Ub.begin(), Ub.end(), U);
        if (W) std::copy(Wb.begin(), Wb.end(), W);
    });
}

int cfg_dns_device_state(cfg_dns* dns, double** u_dev, double** q_dev, int* nkx, int* nkz) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        const StateRef s = dns->it->state();
        if (u_dev) *u_dev = s.u;
        if (q_dev) *q_dev = s.q;
        if (nkx) *nkx = dns->ex->geo->g.nkx;
        if (nkz) *nkz = dns->ex->geo->g.nkz;
    });
}

int cfg_dns_set_graphs(cfg_dns* dns, int enable) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->ex->set_graphs(enable != 0);
    });
}

int cfg_dns_profile(cfg_dns* dns, int enable) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->ex->kt.reset();
        dns->ex->kt.on = enable != 0;
    });
}

int cfg_dns_kernel_times(cfg_dns* dns, double* total_ms, long long* counts) {
    return guarded([&] {
        if (!dns) throw Error(CFG_ERR_ARG, "null dns");
        dns->ex->kt.flush();
        for (int i = 0; i < K_COUNT; ++i) {
            if (total_ms) total_ms[i] = dns->ex->kt.total_ms[i];
            if (counts) counts[i] = dns->ex->kt.count[i];
        }
    });
}

const char* cfg_kernel_name(int id) {
    static const char* names[K_COUNT] = {"k_yinv", "k_xinv", "k_zpass", "k_xfwd", "k_yfwd",
                                         "k_combine", "k_tau_solve", "k_linear", "k_tau_setup"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "";
}

int cfg_nonlinear(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, double* f_spec) {
    return guarded([&] {
        if (!grid || !flags || !u_spec || !f_spec) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        Exec ex(*grid, flags->dealiasing == CFG_DealiasXZ);
        const Geom& g = ex.geo->g;
        PBuf u(ex.pool, ex.nfield), f(ex.pool, ex.nfield);
        std::vector<double> up = upload_packed(g, 3, u_spec, u.p, ex.st);
        auto model = std::make_shared<FlowModel>(*ex.geo, *flags, up, nullptr, nullptr);
        NSECtx nse(ex, model, *flags);
        TickOps ops;
        nse.nonlinear(u.p, f.p, ops);
        ex.run({&ops});
        download_packed(g, 3, f.p, ex.nfield, f_spec, ex.st);
    });
}

// PressureSolver(u, Ubase, Wbase, nu, Vsuck, nonlinearity).solve(q, u)
// (poissonsolver.cpp:270-431): the nonlinear term of u + U (rotation off, as
// the solver's own DNSFlags have it), then the per-mode Poisson solve and wall
// correction. The pressure factors are those of any tau set on this grid.
int cfg_pressure(const cfg_grid* grid, const cfg_flags* flags, const double* Ubase, const double* Wbase,
                 const double* u_spec, double* q_spec) {
    return guarded([&] {
        if (!grid || !flags || !u_spec || !q_spec) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        cfg_flags fl = *flags;
        fl.rotation = 0.0;
        Exec ex(*grid, fl.dealiasing == CFG_DealiasXZ);
        const Geom& g = ex.geo->g;
        PBuf u(ex.pool, ex.nfield), f(ex.pool, ex.nfield), q(ex.pool, ex.nscal);
        std::vector<double> up = upload_packed(g, 3, u_spec, u.p, ex.st);
        auto model = std::make_shared<FlowModel>(*ex.geo, fl, up, Ubase, Wbase);
        NSECtx nse(ex, model, fl);
        TickOps ops;
        nse.nonlinear(u.p, f.p, ops);
        ex.run({&ops});
        auto tau = tau_for(*ex.geo, fl.nu, 1.0, true, ex.st, ex.pool, nullptr);
        if (tau->ready) HIPCHK(hipStreamWaitEvent(ex.st, tau->ready, 0));
        launch_pressure(g, ex.geo->ymats(), ex.geo->HB.p, PressArgs{f.p, u.p, q.p, tau->set, fl.nu}, ex.st);
        HIPCHK(hipGetLastError());
        download_packed(g, 1, q.p, ex.nscal, q_spec, ex.st);
    });
}

int cfg_solve(const cfg_grid* grid, const cfg_flags* flags, double lambda_t, const double* u0_spec,
              const double* rhs_spec, double* u_spec, double* q_spec, double* dPdx) {
    return guarded([&] {
        if (!grid || !flags || !u0_spec || !rhs_spec || !u_spec || !q_spec) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        Exec ex(*grid, flags->dealiasing == CFG_DealiasXZ);
        const Geom& g = ex.geo->g;
        PBuf u(ex.pool, ex.nfield), q(ex.pool, ex.nscal), rhs(ex.pool, ex.nfield);
        std::vector<double> up = upload_packed(g, 3, u0_spec, u.p, ex.st);
        q.zero(ex.st);
        auto model = std::make_shared<FlowModel>(*ex.geo, *flags, up, nullptr, nullptr);
        NSECtx nse(ex, model, *flags);
        nse.reset_lambda({lambda_t});
        (void)upload_packed(g, 3, rhs_spec, rhs.p, ex.st);
        TickOps ops;
        nse.solve(u.p, q.p, {{1.0, rhs.p}}, 0, ops);
        ex.run({&ops});
        // live modes overwritten, the rest of the caller's buffers untouched
        std::vector<double> hu(ex.nfield), hq(ex.nscal);
        HIPCHK(hipMemcpyAsync(hu.data(), u.p, hu.size() * sizeof(double), hipMemcpyDeviceToHost, ex.st));
        HIPCHK(hipMemcpyAsync(hq.data(), q.p, hq.size() * sizeof(double), hipMemcpyDeviceToHost, ex.st));
        HIPCHK(hipStreamSynchronize(ex.st));
        const int Mz = g.Nz / 2 + 1;
        for (int comp = 0; comp < 4; ++comp) {
            const std::vector<double>& src = comp < 3 ? hu : hq;
            double* dst = comp < 3 ? u_spec : q_spec;
            const int cc = comp < 3 ? comp : 0;
            for (int ny = 0; ny < g.Ny; ++ny)
                for (int ikz = 0; ikz < g.nkz; ++ikz)
                    for (int ikx = 0; ikx < g.nkx; ++ikx) {
                        const int kx = ikx <= g.kxd ? ikx : ikx - g.nkx;
                        const int mx = kx >= 0 ? kx : kx + g.Nx;
                        const size_t d = (size_t)ikz + Mz * ((size_t)mx + g.Nx * ((size_t)ny + g.Ny * cc));
                        const size_t s = (((size_t)cc * g.Ny + ny) * g.nkz + ikz) * g.nkx + ikx;
                        dst[2 * d] = src[2 * s];
                        dst[2 * d + 1] = src[2 * s + 1];
                    }
        }
        if (dPdx) *dPdx = nse.dPdx();
    });
}

}  // extern "C"
