/* cfgpu -- C ABI of the MI355X-native channelflow DNS time step.
 *
 * This is the drop-in boundary for channelflow's DNS::advance path. Host code
 * keeps channelflow's FlowField / DNS / DNSFlags interface and calls the GPU
 * through these entry points only: plain pointers and sizes, integer status
 * codes, no exit() inside the library.
 *
 * Spectral field buffers passed across this boundary use channelflow's serial
 * complex layout (channelflow/flowfield.h:387-401), interleaved re/im doubles:
 *     element (mx, my, mz, i) at complex index  mz + (Nz/2+1)*(mx + Nx*(my + Ny*i))
 * with kx = mx for mx <= Nx/2, mx - Nx otherwise, and kz = mz.
 * Modes outside the live set (|kx| > kxd, kz > kzd, and the Nyquist modes) are
 * ignored on input and returned as zero.
 *
 * Status codes: 0 success, <0 failure; cfg_last_error() describes the last
 * failure on the calling thread.
 *
 * Threading: the library keeps process-wide caches (device constants of a
 * flow model, raised LDS limits, pinned staging buffers), so calls must not
 * overlap: use it from one host thread at a time. channelflow's DNS is
 * single-threaded and not re-entrant as well (flowfield.cpp:578-580).
 */
#ifndef CFGPU_H
#define CFGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFG_OK 0
#define CFG_ERR_ARG -1       /* invalid argument / unsupported option     */
#define CFG_ERR_DEVICE -2    /* HIP runtime failure                        */
#define CFG_ERR_NONFINITE -3 /* non-finite value detected in the state     */
#define CFG_ERR_STATE -4     /* call not valid in the object's state       */

/* Enum values follow channelflow/dnsflags.h:26-38. */
enum cfg_baseflow { CFG_ZeroBase = 0, CFG_LinearBase, CFG_ParabolicBase, CFG_LaminarBase, CFG_SuctionBase, CFG_ArbitraryBase };
enum cfg_constraint { CFG_PressureGradient = 0, CFG_BulkVelocity };
enum cfg_stepmethod { CFG_CNFE1 = 0, CFG_CNAB2, CFG_CNRK2, CFG_SMRK2, CFG_SBDF1, CFG_SBDF2, CFG_SBDF3, CFG_SBDF4 };
enum cfg_nonlinearity { CFG_Rotational = 0, CFG_Convection, CFG_Divergence, CFG_SkewSymmetric, CFG_Alternating, CFG_Alternating_ };
enum cfg_dealiasing { CFG_NoDealiasing = 0, CFG_DealiasXZ, CFG_DealiasY, CFG_DealiasXYZ };

/* The DNSFlags fields read by DNS::advance (channelflow/dnsflags.h:106-140). */
typedef struct cfg_flags {
    int baseflow, constraint, timestepping, initstepping, nonlinearity, dealiasing, taucorrection, reserved;
    double nu, Vsuck, rotation, dPdx, dPdz, Ubulk, Wbulk;
    double ulowerwall, uupperwall, wlowerwall, wupperwall, t0, dt;
} cfg_flags;

/* Grid of a FlowField(Nx, Ny, Nz, Nd, Lx, Lz, a, b) (channelflow/flowfield.h:56). */
typedef struct cfg_grid {
    int Nx, Ny, Nz, reserved;
    double Lx, Lz, a, b;
} cfg_grid;

typedef struct cfg_dns cfg_dns;

const char* cfg_last_error(void);
int cfg_device_count(int* n);

/* DNS(fields, flags) / DNS(fields, {Ubase, Wbase}, flags) -- channelflow/dns.h:34-35,
 * dns.cpp:22-62. u_spec: 3-component velocity, q_spec: 1-component pressure
 * (may be NULL for zero). Ubase/Wbase: Chebyshev coefficients (length Ny) of an
 * arbitrary base flow, or NULL to build it from flags.baseflow. */
int cfg_dns_create(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, const double* q_spec,
                   const double* Ubase, const double* Wbase, cfg_dns** out);
int cfg_dns_destroy(cfg_dns* dns);

/* DNS::advance(fields, nSteps) -- dns.cpp:135-163. The velocity and pressure
 * stay resident on the device between calls. Returns CFG_ERR_NONFINITE when
 * the velocity holds a NaN or an infinity afterwards (the test
 * channelflow/cfdsi.cpp:769-774 makes on f^T(u)); the state is kept. */
int cfg_dns_advance(cfg_dns* dns, int nsteps);

/* The caller's `fields` of DNS::advance: set replaces u, q as a caller
 * would between advance() calls; get copies them out. */
int cfg_dns_set_state(cfg_dns* dns, const double* u_spec, const double* q_spec);
int cfg_dns_get_state(cfg_dns* dns, double* u_spec, double* q_spec);

/* DNS::reset_dt / reset_time / reset_gradp / reset_bulkv (dns.cpp:180-232). */
int cfg_dns_reset_dt(cfg_dns* dns, double dt);
int cfg_dns_reset_time(cfg_dns* dns, double t);
int cfg_dns_reset_gradp(cfg_dns* dns, double dPdx, double dPdz);
int cfg_dns_reset_bulkv(cfg_dns* dns, double Ubulk, double Wbulk);

/* DNS::time, dPdx, Ubulk, dt, CFL(u) (dns.cpp:263-330, dnsalgo.cpp:63-67). */
int cfg_dns_time(const cfg_dns* dns, double* t);
int cfg_dns_dpdx(cfg_dns* dns, double* dPdx);
int cfg_dns_ubulk(const cfg_dns* dns, double* Ubulk);
int cfg_dns_cfl(cfg_dns* dns, double* cfl);
/* Ubase(), Wbase() (dns.h:77-78): Chebyshev coefficients, length Ny. */
int cfg_dns_base(const cfg_dns* dns, double* Ubase, double* Wbase);

/* Device pointers to the resident velocity and pressure, valid until the next
 * advance. Compact live-mode layout: complex element (c, ny, ikz, ikx) at
 * ((c*Ny + ny)*nkz + ikz)*nkx + ikx, kx = ikx for ikx <= (nkx-1)/2 and
 * ikx - nkx otherwise, kz = ikz. */
int cfg_dns_device_state(cfg_dns* dns, double** u_dev, double** q_dev, int* nkx, int* nkz);

/* Per-kernel timing with HIP events on the DNS's stream (no counterpart in
 * the reference; used by the benchmark). cfg_dns_profile(dns, 1) clears and
 * enables the counters; cfg_dns_kernel_times fills total milliseconds and
 * launch counts for the CFG_NKERNELS kernel classes named by cfg_kernel_name. */
#define CFG_NKERNELS 9
int cfg_dns_profile(cfg_dns* dns, int enable);
/* hipGraph capture of repeating steps (no counterpart in the reference): 1
 * captures a step whose launch sequence was seen before on its second
 * occurrence and replays it afterwards, 0 (the default) launches eagerly.
 * Results are bit-identical either way; profiling switches replay off while it
 * is on. For one field eager launches are as fast or faster: the SBDF3 history
 * rotation gives three step graphs to capture, and replaying a graph of six
 * small kernels saves no time over launching them (32x33x32: 0.058 ms per step
 * eager, 0.062 replayed). */
int cfg_dns_set_graphs(cfg_dns* dns, int enable);
int cfg_dns_kernel_times(cfg_dns* dns, double* total_ms, long long* counts);
const char* cfg_kernel_name(int id);

/* One NSE::nonlinear evaluation (nse.cpp:383-391): base flow added,
 * nonlinearity per flags.nonlinearity, aliased modes of f zero. */
int cfg_nonlinear(const cfg_grid* grid, const cfg_flags* flags, const double* u_spec, double* f_spec);

/* One NSE::solve with TauSolvers built for lambda_t (nse.cpp:479-575,
 * tausolver.cpp:347-450): rhs -> (u, q) on live modes. u_spec/q_spec are
 * overwritten on live modes only; dPdx (may be NULL) receives the bulk-
 * velocity pressure gradient when flags.constraint is BulkVelocity. */
int cfg_solve(const cfg_grid* grid, const cfg_flags* flags, double lambda_t, const double* u0_spec,
              const double* rhs_spec, double* u_spec, double* q_spec, double* dPdx);

/* One NSE::nonlinear evaluation, as cfg_nonlinear, on device-resident fields,
 * allocating nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc, set-up around HIP graph capture).
 *   u_dev  velocity, f_dev  nonlinear term out: device pointers to 3-component
 *          fields in the compact live-mode layout of cfg_dns_device_state
 *          (complex element (c, ny, ikz, ikx) at ((c*Ny + ny)*nkz + ikz)*nkx + ikx,
 *          nkx = 2 kxd + 1, nkz = kzd + 1, kxd/kzd = Nx/3-1, Nz/3-1 with
 *          DealiasXZ, Nx/2-1, Nz/2-1 without).
 *   stream a hipStream_t, or NULL for the null stream; the call is ordered
 *          after the caller's work on it and returns once f_dev is written.
 * Work buffers and tables live in one fixed device array of
 * CFG_NOALLOC_SCRATCH_BYTES per device (held until the process ends). A grid
 * that needs more (64x65x64 fits, 96x97x96 does not) is rejected with
 * CFG_ERR_ARG. Calls on one device share that array: calls that overlap on one
 * device race. */
#define CFG_NOALLOC_SCRATCH_BYTES (48u << 20)
int cfg_nonlinear_noalloc(const cfg_grid* grid, const cfg_flags* flags, const double* u_dev, double* f_dev,
                          void* stream);

/* PressureSolver(u, Ubase, Wbase, nu, Vsuck, nonlinearity).solve(q, u)
 * (poissonsolver.h:63-97, poissonsolver.cpp:270-431): q with lapl q = -div N(u+U)
 * and dq/dy = nu v'' at the walls, the initial pressure simulateflow.cpp:104-105
 * computes before DNS::advance. nu, Vsuck and nonlinearity come from flags;
 * Ubase/Wbase are Chebyshev coefficients of length Ny, or NULL for the base flow a
 * DNS with these flags builds. u and q in the reference layout. */
int cfg_pressure(const cfg_grid* grid, const cfg_flags* flags, const double* Ubase, const double* Wbase,
                 const double* u_spec, double* q_spec);

/* ---- TauSolver -----------------------------------------------------------------
 * channelflow's per-mode Chebyshev-tau Stokes solver (channelflow/tausolver.h:33-66)
 * for callers that build TauSolvers themselves: it solves
 *     nu u'' - lambda u - grad P = -R,  div u = 0,  u(a) = u(b) = 0
 * for the (kx, kz) Fourier mode of (u, v, w, P). Profiles are ComplexChebyCoeff of
 * length Ny (spectral), passed as Ny interleaved re/im doubles. The tau data and
 * the solves live on the device (the DNS batches all modes in its own kernels;
 * this object serves one mode per call). */
typedef struct cfg_tausolver cfg_tausolver;

/* TauSolver(kx, kz, Lx, Lz, a, b, lambda, nu, Ny, tauCorrection) -- tausolver.cpp:81-176.
 * Ny odd, >= 5. */
int cfg_tausolver_create(int kx, int kz, double Lx, double Lz, double a, double b, double lambda, double nu, int Ny,
                         int taucorrection, cfg_tausolver** out);
int cfg_tausolver_destroy(cfg_tausolver* s);
/* solve(u, v, w, P, Rx, Ry, Rz) -- tausolver.cpp:347-402 */
int cfg_tausolver_solve(cfg_tausolver* s, const double* Rx, const double* Ry, const double* Rz, double* u, double* v,
                        double* w, double* P);
/* solve(u, v, w, P, dPdx, dPdz, Rx, Ry, Rz, umean, wmean) -- tausolver.cpp:404-450: the
 * mean pressure gradients become unknowns and mean(u), mean(w) are imposed; kx = kz = 0
 * only (CFG_ERR_ARG otherwise). */
int cfg_tausolver_solve_bulk(cfg_tausolver* s, const double* Rx, const double* Ry, const double* Rz, double umean,
                             double wmean, double* u, double* v, double* w, double* P, double* dPdx, double* dPdz);
/* verify(u, v, w, P, Rx, Ry, Rz) -- tausolver.cpp:452-629: the sum of the L2 residuals of
 * the momentum, pressure and continuity equations, the wall values, and the squared mean
 * errors (the reference's own sum, including its evaluation of u where it names w at
 * tausolver.cpp:599-600). */
int cfg_tausolver_verify(cfg_tausolver* s, const double* u, const double* v, const double* w, const double* P,
                         const double* Rx, const double* Ry, const double* Rz, double* error);
/* verify(u, v, w, P, dPdx, dPdz, Rx, Ry, Rz, umean, wmean) -- tausolver.cpp:462-629 */
int cfg_tausolver_verify_bulk(cfg_tausolver* s, const double* u, const double* v, const double* w, const double* P,
                              double dPdx, double dPdz, const double* Rx, const double* Ry, const double* Rz,
                              double umean, double wmean, double* error);

/* ---- Batched integrations ---------------------------------------------------
 * Many independent DNS integrations on one device, of any grids. channelflow's
 * callers run them one after another: edge-tracking bisection members
 * (programs/edgetracking.cpp:472-530), multi-shooting segments
 * (channelflow/multiShootingDSI.cpp:70-86), DSI Jacobian probes
 * (channelflow/cfdsi.cpp:678-775). Members on the same grid advance together
 * in batched launches; each member is a full DNS as above (own flags, state,
 * time, init steps). A member's result does not depend on the others. */
typedef struct cfg_batch cfg_batch;

int cfg_batch_create(cfg_batch** out);
int cfg_batch_destroy(cfg_batch* b);
/* New member from spectral u (and q, may be NULL) in the reference layout, as
 * cfg_dns_create; *member receives its id. */
int cfg_batch_add(cfg_batch* b, const cfg_grid* grid, const cfg_flags* flags, const double* u_spec,
                  const double* q_spec, int* member);
int cfg_batch_remove(cfg_batch* b, int member);
/* DNS::advance(fields, nsteps) for every member. Returns CFG_ERR_NONFINITE,
 * naming the members in cfg_last_error(), when a member's velocity holds a NaN
 * or an infinity afterwards (channelflow/cfdsi.cpp:769-774); every member has
 * advanced and keeps its state. */
int cfg_batch_advance(cfg_batch* b, int nsteps);
int cfg_batch_size(const cfg_batch* b, int* nmembers, int* nclasses);
int cfg_batch_get_state(cfg_batch* b, int member, double* u_spec, double* q_spec);
int cfg_batch_set_state(cfg_batch* b, int member, const double* u_spec, const double* q_spec);
int cfg_batch_time(cfg_batch* b, int member, double* t);
/* Whether the member's velocity was finite after its last advance (1) or held
 * a NaN or an infinity (0): the members a CFG_ERR_NONFINITE of
 * cfg_batch_advance names. No counterpart in the reference, whose batch
 * client tests L2Norm(u) itself (channelflow/cfdsi.cpp:769-774). */
int cfg_batch_finite(cfg_batch* b, int member, int* finite);
int cfg_batch_dpdx(cfg_batch* b, int member, double* dPdx);
/* DNS::CFL(u) of one member's current velocity (dns.cpp:282-289). */
int cfg_batch_cfl(cfg_batch* b, int member, double* cfl);
/* DNS::reset_dt(dt) of one member (dns.cpp:180-185, dnsalgo.cpp:191-200): its
 * tau data are rebuilt for dt and its next advance restarts the
 * initialisation steps, as the reference's callers do each dT
 * (channelflow/cfdsi.cpp:733-767, programs/edgetracking.cpp:517-520). */
int cfg_batch_reset_dt(cfg_batch* b, int member, double dt);
/* DNS::advance(fields, nsteps[k]) of the `count` listed members, together:
 * members whose dt differ advance their own step counts over one dT. Members
 * not listed do not move. Non-finite states as cfg_batch_advance. */
int cfg_batch_advance_members(cfg_batch* b, int count, const int* members, const int* nsteps);
/* Kernel timing summed over the batch's streams (as cfg_dns_profile). */
int cfg_batch_profile(cfg_batch* b, int enable);
int cfg_batch_kernel_times(cfg_batch* b, double* total_ms, long long* counts);
/* hipGraph capture of repeating ticks (no counterpart in the reference).
 * mode 0: eager launches; 1 (the default): one graph per grid class and
 * tick, each class on its own stream, the classes running concurrently;
 * 2: one graph per tick for the whole batch, every grid class a parallel
 * branch of it (one launch per tick, but ROCm's launch of such a graph
 * costs more host time than the per-class launches together, with the same
 * GPU time). A tick whose launch sequence was seen before is captured on its
 * second occurrence and replayed afterwards; results are bit-identical to
 * eager launches in every mode. Profiling switches replay off while it is
 * on. */
int cfg_batch_set_graphs(cfg_batch* b, int mode);
int cfg_batch_graph_stats(const cfg_batch* b, long long* captures, long long* replays);
/* Ticks run so far (a tick is one nonlinear evaluation of every member below
 * its target), graph launches among them, and host microseconds spent in them
 * (planning, launch tables, launches). No counterpart in the reference. */
int cfg_batch_tick_stats(const cfg_batch* b, long long* ticks, long long* graph_launches, double* host_us);

/* ---- Serving ----------------------------------------------------------------
 * Continuous batching of integration requests (no counterpart in the
 * reference). A request is (grid, flags, initial u and q, horizon in steps).
 * Queued requests are admitted in order while their device memory fits the
 * budget; a request that does not fit yet is passed by later ones that do
 * (backfilling) until it has waited 256 server steps, after which nothing
 * passes it; every admitted request advances one tick (one nonlinear evaluation)
 * per cfg_server_step together with the others; a request that reaches its
 * horizon is retired at once and its result is copied back without stalling
 * the rest. */
typedef struct cfg_server cfg_server;

int cfg_server_create(double budget_bytes, cfg_server** out);
int cfg_server_destroy(cfg_server* s);
int cfg_server_submit(cfg_server* s, const cfg_grid* grid, const cfg_flags* flags, const double* u_spec,
                      const double* q_spec, int nsteps, long long* request);
/* One server step. done receives up to max_done ids of requests whose results
 * reached host memory since the last call (*ndone of them). */
int cfg_server_step(cfg_server* s, long long* done, int max_done, int* ndone, int* nrunning, int* nqueued);
/* Final state (reference layout) and time of a completed request; releases it.
 * Returns CFG_ERR_NONFINITE when that velocity holds a NaN or an infinity
 * (channelflow/cfdsi.cpp:769-774); the outputs are written all the same. */
int cfg_server_result(cfg_server* s, long long request, double* u_spec, double* q_spec, double* t);
/* Device memory the server holds for its integrations, which admission keeps
 * within the budget: pool blocks (in use or free, at hipMalloc's rounding),
 * work buffers and tau data. A grid class's fixed tables (launch-table arena,
 * geometry) are outside the budget; a class that empties hands its memory back. */
int cfg_server_memory(const cfg_server* s, double* used_bytes);

#ifdef __cplusplus
}
#endif
#endif /* CFGPU_H */
