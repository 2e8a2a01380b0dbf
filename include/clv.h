/*
 * clv.h -- C ABI of the MI355X complex-Langevin engine (libclv.so).
 *
 * The reference (DoAmann/CL_var_dt) has no plugin or FFI boundary: its host
 * driver calls C++ classes in one translation unit.  Each entry point below
 * replaces one piece of that class surface; the reference member it stands in
 * for is cited beside it.  Plain C, no exceptions across the boundary, status
 * is an int (CLV_OK == 0, negative on error), the message of the last failure
 * is returned by clv_last_error().
 *
 * Field layout at the boundary is the reference's device layout
 * (Lattice.cu:2-5): complex128 interleaved re/im, index
 *   ((((c*Nt + tau)*X + x)*Y + y)*Z + z),  c = component 0..3.
 * Internally the library keeps its own layout (chain-major, FP64 or FP32).
 *
 * Threading: a context owns its device memory and one HIP stream.  Calls on
 * one context must not overlap; separate contexts may share a device.
 * clv_propagate, clv_propagate_plain's device work and clv_evaluate are
 * asynchronous to the host; getters
 * synchronize the context's stream.  The library never prints.
 */
#ifndef CLV_H_
#define CLV_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLV_OK 0
#define CLV_ERR_ARG (-1)         /* bad argument / unsupported shape        */
#define CLV_ERR_HIP (-2)         /* HIP runtime error                       */
#define CLV_ERR_NOMEM (-3)       /* device allocation failed                */
#define CLV_ERR_BUDGET (-4)      /* request does not fit the HBM budget     */
#define CLV_ERR_STATE (-5)       /* call not valid in the current state     */
#define CLV_ERR_NOTFOUND (-6)    /* unknown request / chain id              */
#define CLV_ERR_CANCELLED (-7)   /* run cancelled by clv_server_cancel      */

#define CLV_FP64 0
#define CLV_FP32 1

/* Conj-drift mode.  In the reference the psi-bar drift's potential term multiplies
 * psi(tau+1) (Propagator_without_cutoff.cu:71,88,105,122).  CLV_CONJ_DRIFT_REFERENCE
 * keeps that; CLV_CONJ_DRIFT_ACTION uses psibar(tau+1) (the action derivative). */
#define CLV_CONJ_DRIFT_REFERENCE 0
#define CLV_CONJ_DRIFT_ACTION 1

/* One chain's lattice: X x Y x Z space, Nt imaginary-time slices, 4 comps.
 * Stands in for the XSIZE/YSIZE/ZSIZE/TAUSIZE macros (Parameters.h:8-11). */
typedef struct clv_shape {
  int32_t nx, ny, nz, nt;
} clv_shape;

/* Per-chain physics and controller parameters.  The reference takes MU and P
 * as run-time arguments of Propagator::propagate (Propagator_without_cutoff.cu:406)
 * and everything else as macros (Parameters.h:13-31); here all are run time. */
typedef struct clv_params {
  double mu;           /* chemical potential MU (params.txt mu)            */
  double P;            /* linear Zeeman P (params.txt P)                    */
  double G;            /* density coupling c0 (Parameters.h:15 G)           */
  double c1;           /* spin coupling; 0 reproduces the reference         */
  double Q;            /* quadratic Zeeman (Parameters.h:24 Q)              */
  double eps;          /* EPS = beta / Nt (Parameters.h:13)                 */
  double a_x;          /* lattice spacing a_x (Parameters.h:14)             */
  double DT;           /* controller base step DT (Parameters.h:29)         */
  double dt_min;       /* lower clamp, 1e-8 (Propagator_without_cutoff.cu:391) */
  double dt_max;       /* upper clamp, 1e-2 (Propagator_without_cutoff.cu:392) */
  double alpha;        /* blend, 0.9 (Propagator_without_cutoff.cu:402)     */
  double factor;       /* drift prefactor, 1 in propagate (:427)            */
  uint64_t seed;       /* Philox key of this chain's noise stream           */
  int32_t kin_dims;    /* 1: p^2 = p_x^2 (reference, :196); 3: px^2+py^2+pz^2 (two-pass
                          schedule; Y, Z in the supported lengths; not in the server) */
  int32_t conj_drift_mode;     /* CLV_CONJ_DRIFT_REFERENCE or CLV_CONJ_DRIFT_ACTION                 */
  int32_t noise_on;    /* 1: propagate (:406); 0: propagate_without_noise (:444) */
  int32_t reserved;
} clv_params;

/* Controller state of one chain after its last step. */
typedef struct clv_step_stats {
  double dt;           /* dt used by the last step                          */
  double max_drift;    /* max |D_0| of the last step                        */
  double mean_drift;   /* sum |D_0| / (4 V_s Nt) of the last step           */
  double t_langevin;   /* sum of dt over all steps                          */
  int64_t steps;       /* steps taken                                       */
  int64_t clamp_low;   /* steps whose dt hit the lower clamp                */
  int64_t clamp_high;  /* steps whose dt hit the upper clamp                */
  int32_t nonfinite;   /* 1 once a non-finite drift has been seen           */
  int32_t reserved;
} clv_step_stats;

/* Running observable means, in the form the reference writes them
 * (Observables_without_cutoff.cu:1527-1547, 1614-1634): sums divided by
 * (Nt * count) and, for the numbers, by V_s as in write_n_tot. */
typedef struct clv_observables {
  int64_t count;       /* number of evaluations (Observables::get_counter)  */
  double n_tot;        /* total particle number (write_n_tot)               */
  double n_comp[4];    /* n_11, n_12, n_21, n_22                            */
  double energy;       /* <H> without the -mu N term (new estimator)        */
  double Fx_tot, Fy_tot, Fz_tot;  /* sum_x of F_a(x)                        */
  double density;      /* n_tot / V_s                                       */
  double Sz;           /* Fz_tot / V_s                                      */
} clv_observables;

typedef struct clv_ctx clv_ctx;

/* ---- library ---------------------------------------------------------- */
const char *clv_version(void);
/* Message of the last failed call on ctx (or of the last failed create when
 * ctx is NULL).  Never NULL. */
const char *clv_last_error(const clv_ctx *ctx);
/* Reference defaults: Parameters.h:13-31 and Propagator_without_cutoff.cu:382-404. */
void clv_default_params(clv_params *p);
/* Propagator::compute_new_dt (Propagator_without_cutoff.cu:382-404), on the
 * host.  clamp: 0 none, -1 low clamp applied, +1 high clamp applied. */
double clv_compute_new_dt(double mean_drift, double max_drift,
                          const clv_params *p, int32_t *clamp);

/* ---- context: replaces `new ComplexLattice` x4 + `new Propagator`
 * (main.cu:94-103, Lattice.cu:108-125, Propagator_without_cutoff.cu:316-343) */
int clv_create(const clv_shape *shapes, int32_t nchains, int32_t precision,
               int32_t device, clv_ctx **out);
int clv_destroy(clv_ctx *ctx);
int clv_num_chains(const clv_ctx *ctx);
/* Bytes of device memory the context holds. */
int64_t clv_device_bytes(const clv_ctx *ctx);

int clv_set_params(clv_ctx *ctx, int32_t chain, const clv_params *p);

/* Fields in the reference layout (see top).  psi / psibar each hold
 * 2*4*V_s*Nt doubles.  Replaces ComplexLattice::set_mean_field / operator=
 * (Lattice.cu:190-214) for initialisation and readback. */
int clv_set_fields(clv_ctx *ctx, int32_t chain, const double *psi,
                   const double *psibar);
int clv_get_fields(clv_ctx *ctx, int32_t chain, double *psi, double *psibar);
/* psi_c = sqrt(n0) + xi_c, xi_c complex Gaussian with variance `var` per
 * real part from the chain's Philox key (init stream); psibar = conj(psi). */
int clv_init_fields(clv_ctx *ctx, int32_t chain, double n0, double var);

/* ---- the per-step update: Propagator::propagate (Propagator_without_cutoff.cu:406-442)
 * Device-resident controller: every chain advances nsteps steps, dt per
 * chain computed on the device from that chain's own drift.  eval_every > 0
 * accumulates observables after step i when (chain's step index i) %
 * eval_every == 0, as main.cu:152-156 does after propagate. */
int clv_propagate(clv_ctx *ctx, int64_t nsteps, int32_t eval_every);

/* Plain method: the reference's host-driven step.  Per step: drift + max/mean
 * reduction on the device, a blocking device->host copy of (max, sum), dt on
 * the host (clv_compute_new_dt), then the update with dt passed by value
 * (Propagator_without_cutoff.cu:427-441).  Single chain (chain index). */
int clv_propagate_plain(clv_ctx *ctx, int32_t chain, int64_t nsteps,
                        int32_t eval_every);

/* Building blocks of the plain step for a host driver that mirrors
 * Propagator::propagate itself:
 *   clv_drift_stats  = calculate_drift_without_noise + get_max_and_mean_drift
 *                      (Propagator_without_cutoff.cu:427-429, 351-380)
 *   clv_apply_update = multiply_by_timestep + add_noise + fields += drift
 *                      with the given dt (Propagator_without_cutoff.cu:432-441) */
int clv_drift_stats(clv_ctx *ctx, int32_t chain, double *max_drift,
                    double *mean_drift);
int clv_apply_update(clv_ctx *ctx, int32_t chain, double dt);

int clv_get_step_stats(clv_ctx *ctx, int32_t chain, clv_step_stats *out);
/* Restore a chain's controller state (steps taken = noise counter, Langevin
 * time, last dt / max / mean, clamp counters) -- with clv_set_fields this
 * resumes a run exactly, since the noise is a pure function of (seed, step,
 * site).  No counterpart in the reference (its save/read helpers,
 * SaveResults.cu:797-932, are never called and do not save the RNG). */
int clv_set_step_stats(clv_ctx *ctx, int32_t chain, const clv_step_stats *in);
int clv_synchronize(clv_ctx *ctx);

/* ---- observables: Observables::evaluate / write_* (Observables_without_cutoff.cu:1155-1524, 1527-1634) */
int clv_evaluate(clv_ctx *ctx, int32_t chain);   /* chain < 0: all chains */
/* Fx, Fy, Fz: V_s doubles each in (x,y,z) order, written form; may be NULL. */
int clv_read_observables(clv_ctx *ctx, int32_t chain, clv_observables *out,
                         double *Fx, double *Fy, double *Fz);
int clv_reset_observables(clv_ctx *ctx, int32_t chain);

/* ---- kernel selection (project autotuner) ------------------------------ */
/* Times the candidate kernel schedules and row-transform splits of every
 * shape bucket of ctx on its current state (state is saved and restored) and
 * keeps the fastest.  When
 * cache_path is non-NULL the choice is read from / appended to that JSON-lines
 * file, keyed by shape, precision and device name. */
int clv_autotune(clv_ctx *ctx, const char *cache_path);
/* Force a schedule: 0 = auto (fused when tiles x chains >= 512 workgroups,
 * else two-pass; no timing), 1 = two-pass (reduce then update),
 * 2 = fused single pass (update + next step's reduction). */
int clv_set_schedule(clv_ctx *ctx, int32_t schedule);
/* The schedule the chain's bucket runs (1 or 2). */
int clv_get_schedule(const clv_ctx *ctx, int32_t chain);
/* Radix split of the row transform of the chain's bucket (every chain of its
 * shape): the length N1 of the first level of the two-level transform of
 * rows of X sites, X = N1 * N2 -- e.g. 128 = 16 x 8 (default) or 8 x 16.  The
 * reference leaves this choice to cufftPlanMany (Lattice.cu:135-156).  Splits
 * regroup the transform's arithmetic, so results differ at the rounding level
 * only (FP64 ~1e-16, FP32 ~1e-7 relative per step); clv_autotune times them.
 * n1 = 0 restores the default; CLV_ERR_ARG for a split this library does not
 * build for X or that does not fit LDS at the shape's row tile.
 * clv_get_row_split returns the bucket's N1. */
int clv_set_row_split(clv_ctx *ctx, int32_t chain, int32_t n1);
int clv_get_row_split(const clv_ctx *ctx, int32_t chain);
/* Options of the fused schedule for every bucket that has it (project
 * extension, no counterpart in the reference).  line_sets: LDS line sets,
 * 1 or 2 (two alternate by tau parity); batch: tau slices whose next-step
 * K_0 row transforms run together, 1, 2 or 4.  0 keeps the current value.
 * Results do not depend on either; a combination that does not fit 64 KB of
 * LDS for some shape is rejected.  clv_autotune times them. */
int clv_set_fused_options(clv_ctx *ctx, int32_t line_sets, int32_t batch);
/* Tau blocks of the fused schedule for every bucket (project extension, no
 * counterpart in the reference): the tau range splits into `blocks` equal
 * blocks, one workgroup per (row tile, chain, block), so few chains still
 * give enough workgroups; the first slice of each block gets its next-step
 * |D_0| partial from a short drift-statistics pass after the march.
 * blocks > 1 must divide Nt into blocks of at least 2 slices; 1 = one block
 * (the whole tau range); 0 = the untimed choice (the fewest blocks of at
 * least 8 slices that give 2048 waves, e.g. 512 four-wave workgroups).
 * Results do not depend on it; clv_autotune times it. */
int clv_set_fused_blocks(clv_ctx *ctx, int32_t blocks);
/* clv_propagate replays captured HIP graphs of its steps (on by default): a
 * cadence period with the observable kernels only where they evaluate when
 * eval_every > 0 and each bucket's chains share a step count, else two steps
 * with the observable kernels gated per chain; off runs every step eagerly.
 * Results are identical. */
int clv_set_graphs(clv_ctx *ctx, int32_t on);
/* Capture and instantiate the step graph that the next clv_propagate with
 * this eval_every would replay (for the buckets' current schedules, state
 * slots and cadence phases), without launching it (a project extension with no
 * counterpart in the reference: for callers that keep the one-time capture out
 * of a timed or latency-critical region).  No-op with graphs off or kernel
 * timing on; a later change of schedule or blocks drops it. */
int clv_prepare_graph(clv_ctx *ctx, int32_t eval_every);

/* ---- kernel timing -----------------------------------------------------
 * With profiling on, every kernel the context launches is bracketed by HIP
 * events on the context's stream; clv_kernel_time returns the accumulated
 * device time (ms) and launch count of one kernel kind:
 * 0 drift_stats, 1 controller, 2 update, 3 obs, 4 obs_finish, 5 fused step,
 * 6 transverse kinetic term (kin_dims = 3).
 * (No counterpart in the reference, which times only the whole run, main.cu:59,159.) */
int clv_set_profiling(clv_ctx *ctx, int32_t on);
int clv_kernel_time(clv_ctx *ctx, int32_t kind, double *total_ms, int64_t *launches);

/* ---- allocation-free step ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: one Langevin step of one chain held in caller-owned device
 * buffers, the step of clv_propagate (device-resident dt, the two-pass kernels,
 * identical results) without observables, calling no hipMalloc or hipFree.
 *   psi, psibar          device, the chain's fields in the library's internal
 *                        layout: complex64 (CLV_FP32) or complex128 (CLV_FP64)
 *                        value (site, c) at site*4 + c, site = (tau*Y*Z + y*Z + z)*X + x;
 *   psi_out, psibar_out  device, the fields after the step (same layout; must
 *                        not overlap the inputs);
 *   stats                host, in: the chain's controller state (steps is the
 *                        noise counter; all zero for a fresh chain); out: the
 *                        state after the step;
 *   stream               hipStream_t to run on (NULL: the default stream); the
 *                        call synchronizes it before returning.
 * The chain's parameter record, controller state, row-transform tables and its
 * per-(tau, row tile) |D_0| partials live in a static 1 MiB device array that
 * each GPU holds until the process ends; a shape whose Nt x row tiles x 16
 * bytes of partials does not fit is rejected with CLV_ERR_ARG (256x256x1x128
 * fits, 256x512x1x256 does not).  Every call on one GPU shares that array:
 * calls that overlap on one GPU race.  kin_dims must be 1.  Errors:
 * clv_last_error(NULL). */
int clv_step_noalloc(const clv_shape *shape, int32_t precision, const clv_params *params,
                     const void *psi, const void *psibar, void *psi_out, void *psibar_out,
                     clv_step_stats *stats, int32_t device, void *stream);

/* ---- serving: continuous batching of Langevin runs -----------------------
 * No counterpart in the reference, which runs one chain per process and
 * batches runs as SLURM job arrays polled from Python (mylib.py:511-543,
 * 257-276).  A server owns a worker thread and device-resident pools of chain
 * slots, one pool per lattice shape, each stepped on its own HIP stream.
 * Requests are admitted into free slots as they arrive, by HBM budget; a run
 * retires exactly after its nsteps steps and its observables are returned.
 * A run's result depends only on its request, never on its slot or on the
 * other runs in flight. */
typedef struct clv_request {
  clv_shape shape;
  clv_params params;
  int64_t nsteps;      /* Langevin steps of the run                          */
  int32_t eval_every;  /* observable cadence; 0: once, after the last step   */
  int32_t reserved;
  int64_t eval_from;   /* first step index that may be evaluated             */
  double n0, var;      /* initial fields (see clv_init_fields)               */
} clv_request;

#define CLV_RUN_QUEUED 0
#define CLV_RUN_RUNNING 1
#define CLV_RUN_DONE 2

typedef struct clv_result {
  int64_t id;
  int32_t status;      /* CLV_RUN_*, or a negative CLV_ERR_* code            */
  int32_t reserved;
  clv_observables obs;
  clv_step_stats stats;
  double t_submit, t_admit, t_done;  /* clv_server_clock() seconds          */
} clv_result;

typedef struct clv_server_info {
  int64_t submitted, completed, running, queued, failed;
  int64_t bytes_in_use, budget_bytes, pools;
  double site_updates;   /* site-updates executed so far                    */
  double busy_seconds;   /* wall time with at least one run in flight        */
  int64_t cancelled;     /* runs ended by clv_server_cancel                  */
  double pool_setup_seconds;  /* worker time spent creating slot pools      */
} clv_server_info;

typedef struct clv_server clv_server;

/* hbm_budget_bytes: device memory the pools may hold (0: 80% of free). */
int clv_server_create(int32_t device, int32_t precision, int64_t hbm_budget_bytes,
                      clv_server **out);
int clv_server_destroy(clv_server *srv);
/* Stamps t_submit on entry, before any lock: time a client spends blocked in
 * the call counts toward the run's latency.  Never waits for device work. */
int clv_server_submit(clv_server *srv, const clv_request *req, int64_t *id);
/* Non-blocking; out->status tells where the run is. */
int clv_server_poll(clv_server *srv, int64_t id, clv_result *out);
/* Blocks until the run is done or timeout_s passes (CLV_ERR_STATE). */
int clv_server_wait(clv_server *srv, int64_t id, double timeout_s, clv_result *out);
/* Forgets a finished (done or failed) run's record, which the server keeps
 * for poll/wait until then; CLV_ERR_STATE while it is queued or running. */
int clv_server_release(clv_server *srv, int64_t id);
/* Cancels a queued or running run.  A queued run leaves the queue; a running
 * run's slot is freed at its pool's next tick boundary (the other slots of the
 * pool are unaffected) and its memory is available to new runs from then on.
 * The run ends with status CLV_ERR_CANCELLED and the step statistics it had
 * reached.  CLV_ERR_STATE if the run has already finished.  The reference has
 * no in-process control of a run: it scans SLURM logs for failed runs and
 * relaunches them as new array jobs (mylib.py:160-199); a client of this
 * server cancels the run and submits it again. */
int clv_server_cancel(clv_server *srv, int64_t id);
int clv_server_info_get(clv_server *srv, clv_server_info *out);
/* The server's monotonic clock, seconds. */
double clv_server_clock(void);
/* Last error message; a copy valid until the calling thread's next call. */
const char *clv_server_error(const clv_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* CLV_H_ */
