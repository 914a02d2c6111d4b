/*
 * gradpower_hip.h -- C ABI of the MI355X batched backward-Euler DAE integrator.
 *
 * Replaces the GPU entry points of GradPower.jl's CUDA extension
 * (ext/GradPowerCUDAExt/GradPowerCUDAExt.jl):
 *   GpuBatchedLayout(dp, ps, M)                 ext:466          -> gp_create + gp_set_state
 *   integrate_gpu_cudss!(gbl, ps, tf; dt, tol)  ext:2924-2996    -> gp_integrate(GP_METHOD_MONO_FP64)
 *   integrate_gpu_schur_cudss!(gbl, ps, tf; ..) ext:3008-3080    -> gp_integrate(GP_METHOD_SCHUR_*)
 *   integrate_gpu_shared_multi!(..., fault_buses, rfaults, ton, toff)
 *                                               ext:2640-2708    -> gp_set_events (per-scenario faults)
 *                                                                   + gp_integrate(GP_METHOD_SCHUR_SHARED)
 *   integrate_gpu_shared!(gbl, ps, tf; ..)      ext:2296-2375    -> gp_integrate(GP_METHOD_SCHUR_SHARED)
 *   TripLineEvent / create_trip_line_event     src/GradPower.jl:140-167,509-570
 *   (CPU integrate! only, src/dynamics.jl:536-540,616-624)        -> gp_set_line_trips (per scenario)
 *   DisconnectDeviceEvent (ZIP loads)          src/GradPower.jl:130-143,681-698
 *   (CPU integrate! only, src/dynamics.jl:533-535,612-615)        -> gp_set_load_drops (per scenario)
 *   trajs[m] = z_hist[:, m, :]                  ext:3075-3079    -> gp_get_trajectory
 *   _residual_all_scenarios_gpu! / _jacobian_all_scenarios_gpu!
 *                                               ext:1318-1516    -> gp_eval_residual / gp_eval_jacobian
 *
 * Conventions
 *   - All indices are 0-based (the Julia wrapper subtracts 1).
 *   - The system descriptor holds borrowed host pointers, read only during
 *     gp_create; the library copies what it needs.
 *   - Every call returns GP_OK (0) or a negative gp_status; gp_last_error()
 *     returns a thread-local message for the last failure on this thread.
 *     No C++ exception crosses the ABI.
 *   - A context owns its device memory and one HIP stream on the device that
 *     was current at gp_create.  Calls on one context must be serialised by
 *     the caller; distinct contexts may be used concurrently.
 */
#ifndef GRADPOWER_HIP_H
#define GRADPOWER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum gp_status {
  GP_OK = 0,
  GP_ERR_BADARG = -1,
  GP_ERR_HIP = -2,
  GP_ERR_OOM = -3,
  GP_ERR_SINGULAR_PIVOT = -4,
  GP_ERR_NEWTON_NOT_CONVERGED = -5,
  GP_ERR_NAN = -6,
  GP_ERR_UNSUPPORTED = -7
};

/* ---- device tables ------------------------------------------------------- */
/* Each table is a row-major int32 array [n_dev[t] x GP_COLS_<t>] of absolute
 * 0-based z indices (-1 = not wired), parameter offsets into p and 0-based bus
 * indices.  They correspond to the reference SoA tables after reorder_state!
 * (src/layout.jl:61-179, src/clusters.jl:339-466). */
enum gp_devtype {
  GP_GENROU = 0,   /* diff_ptr, alg_ptr, par_ptr, bus, pm_idx, efd_idx        */
  GP_IEESGO = 1,   /* diff_ptr, alg_ptr, par_ptr, w_idx                        */
  GP_TGOV1 = 2,    /* diff_ptr, alg_ptr, par_ptr, w_idx                        */
  GP_SEXS = 3,     /* diff_ptr, par_ptr, vr_idx, vs_idx                        */
  GP_ESDC1A = 4,   /* diff_ptr, par_ptr, vr_idx, vs_idx                        */
  GP_IEEEST = 5,   /* diff_ptr, alg_ptr, par_ptr, w_idx                        */
  GP_ZIPLOAD = 6,  /* bus, par_ptr                                             */
  GP_STATICGEN = 7,/* bus, alg_ptr (-1 for PQ), par_ptr, bus_type (1 PQ,2 PV,3 slack) */
  GP_NDEVTYPES = 8
};
#define GP_COLS_GENROU 6
#define GP_COLS_IEESGO 4
#define GP_COLS_TGOV1 4
#define GP_COLS_SEXS 4
#define GP_COLS_ESDC1A 4
#define GP_COLS_IEEEST 4
#define GP_COLS_ZIPLOAD 2
#define GP_COLS_STATICGEN 4

/* Cluster table row (src/clusters.jl:16-28): a generator with its controllers,
 * occupying z[w_start .. w_start+w_size).  Controller/generator fields are
 * row indices into the device tables (-1 = absent). */
enum gp_cluster_col {
  GP_CL_WSTART = 0, GP_CL_WSIZE, GP_CL_BUS, GP_CL_TRIVIAL, GP_CL_GENROU,
  GP_CL_GOVTYPE,   /* 0 none, GP_IEESGO or GP_TGOV1 */
  GP_CL_GOV, GP_CL_EXCTYPE /* 0 none, GP_SEXS or GP_ESDC1A */, GP_CL_EXC,
  GP_CL_PSS, GP_CL_STATICGEN, GP_CLUSTER_COLS
};

typedef struct gp_system_desc {
  int32_t n;          /* n_sys = n_diff + n_alg + 2*nbus                      */
  int32_t nbus;
  int32_t net_ptr;    /* z index of vr of bus 0; vr_b = net_ptr+2b, vi_b = +1  */
  int32_t n_par;
  const uint8_t *is_diff;      /* [n] 1 = differential state                   */
  /* realified network admittance (2nbus x 2nbus), CSR (src/network.jl:87)   */
  int32_t y_nnz;
  const int32_t *y_rowptr;     /* [2*nbus+1] */
  const int32_t *y_colind;     /* [y_nnz]    */
  const double *y_val;         /* [y_nnz]    */
  int32_t n_dev[GP_NDEVTYPES];
  const int32_t *dev[GP_NDEVTYPES];
  const double *genrou_u0;     /* [n_genrou*2] e_fd, p_m used when unwired   */
  int32_t n_clusters;
  const int32_t *clusters;     /* [n_clusters*GP_CLUSTER_COLS]              */
} gp_system_desc;

/* One bus fault of one scenario (src/GradPower.jl:117-128, ext:2640-2650):
 * conductance 1/r_fault on the bus's vr/vi rows from the end of step
 * step_on until the end of step step_off (steps counted from 1; both >= 1). */
typedef struct gp_event {
  int32_t scenario;
  int32_t bus;
  double r_fault;
  int32_t step_on;
  int32_t step_off;
} gp_event;

/* One line trip of one scenario (TripLineEvent, src/GradPower.jl:140-167):
 * the line's admittance stamp y -- what the line adds to Y at (from, from),
 * (from, to), (to, from), (to, to), each as re, im (create_trip_line_event,
 * src/GradPower.jl:509-542) -- is taken out of the network at the end of step
 * step_trip and put back at the end of step step_reclose (0: never), with a
 * dt = 0 re-solve after each, as a fault's activation and clearing
 * (src/dynamics.jl:536-540,599-645).  Both off-diagonal blocks must be in
 * the pattern of Y. */
typedef struct gp_line_trip {
  int32_t scenario;
  int32_t from_bus;
  int32_t to_bus;
  int32_t step_trip;
  int32_t step_reclose;
  int32_t reserved;
  double y[8];                 /* yff, yft, ytf, ytt as (re, im) pairs */
} gp_line_trip;

/* One ZIP load of one scenario disconnected (DisconnectDeviceEvent,
 * src/GradPower.jl:130-143,681-698) at the end of step `step`, for good, with
 * a dt = 0 re-solve after it: its current leaves the bus's rows (the
 * reference skips offline devices in its kernels).  `load` is the row of the
 * ZIP load in desc->dev[GP_ZIPLOAD].  reserved: 0. */
typedef struct gp_load_drop {
  int32_t scenario;
  int32_t load;
  int32_t step;
  int32_t reserved;
} gp_load_drop;

/* One device of one scenario taken offline (DisconnectDeviceEvent) at the
 * end of step `step`, for good, with a dt = 0 re-solve after it.  The
 * reference's kernels skip an offline device into a zeroed F
 * (ext:1340,1072-1199; the `online` checks of src/kernels/), so:
 *   GP_ZIPLOAD: row `row` of desc->dev[GP_ZIPLOAD]; its current leaves the
 *     bus's rows (as gp_load_drop);
 *   GP_SEXS, GP_ESDC1A: row `row` of that exciter table; its rows of F are
 *     zero, so backward Euler holds its states (and the generator's field
 *     voltage) at their values when it went offline.
 * Other types are refused (GP_ERR_UNSUPPORTED): a device with algebraic
 * states (GENROU, governors, IEEEST, static generators) whose rows leave F
 * and J leaves those states undetermined (a singular Jacobian). */
typedef struct gp_device_drop {
  int32_t scenario;
  int32_t type;      /* gp_device_type */
  int32_t row;
  int32_t step;
} gp_device_drop;

enum gp_method {
  GP_METHOD_MONO_FP64 = 0,    /* monolithic Newton, sparse LU of J refactored every iteration */
  GP_METHOD_SCHUR_FP64 = 1,   /* cluster Schur reduction, FP64 factors                        */
  GP_METHOD_SCHUR_FP32 = 2,   /* cluster Schur reduction, FP32 block/network factors,
                                 FP64 residual, update and states (inexact Newton)           */
  GP_METHOD_SCHUR_SHARED = 3  /* cluster Schur reduction with ONE FP64 factorisation of the
                                 reduced network shared by every scenario: P = S of
                                 scenario 0 at its state when gp_integrate starts, step dt,
                                 no event in force; each scenario's faults enter its solves
                                 exactly as a Woodbury correction of rank two per faulted
                                 bus (ext:2135-2290, 2409-2495, 2523-2708).  Cluster blocks
                                 are exact at every iteration; the network part is frozen
                                 (inexact Newton).  A solve that fails on P (itermax, or
                                 retry_window corrections without halving |F|inf) runs
                                 again from the step's start as exact FP64 Schur Newton
                                 (gp_stats.exact_retries), where the reference keeps the
                                 unconverged state.  Not with line trips, load drops,
                                 lock_step, gp_adjoint, or reduced networks whose solution
                                 does not fit in LDS (GP_ERR_UNSUPPORTED).                 */
};

enum gp_refactor_policy {
  GP_REFACTOR_EVERY_ITER = 0,  /* numeric factorisation at every Newton iteration (the
                                  reference's cuDSS factorization/refactorization cycle) */
  GP_REFACTOR_FIRST_ITER = 1,  /* factorise at the first iteration of each Newton solve
                                  and reuse it for the later iterations (Schur methods)  */
  GP_REFACTOR_ADAPTIVE = 2,    /* keep the factorisation across time steps; factorise
                                  again after a fault event or a change of dt, and
                                  whenever an iteration fails to reduce |F|inf by
                                  refactor_theta (default 30x)                         */
  GP_REFACTOR_ASYNC = 3        /* the iterates of ADAPTIVE, but a refactorisation asked
                                  for after the first iteration of a solve runs on a
                                  second stream: the scenario waits for it while the
                                  other scenarios keep iterating, instead of every
                                  scenario waiting in the lock step                   */
};

typedef struct gp_options {
  int32_t record_all;      /* 1: record every state each step; 0: only record_idx  */
  int32_t n_record;        /* number of recorded z indices when record_all == 0     */
  const int32_t *record_idx;
  int32_t refactor_policy; /* gp_refactor_policy for the Schur methods              */
  int32_t stall_window;    /* 0 (default, the reference's behaviour): a Newton solve
                              runs to itermax.  k > 0: a scenario whose |F|inf has
                              not halved within its last k iterations, or exceeds
                              1e10, or is not finite, stops at once and is reported
                              as not converged (failed contingencies cost k rounds
                              of the lock step instead of itermax)               */
  int32_t profile;         /* k >= 1: time every k-th launch of the numeric
                              factorisation and of the batched triangular solve
                              with HIP events on the stream it runs on
                              (gp_stats kernel fields cover the timed
                              launches); 0: no timing                           */
  int32_t lock_step;       /* 0 (default): every scenario advances on its own --
                              a scenario whose step converged moves on while the
                              others iterate; 1: all scenarios take step k
                              together (a step lasts as long as its slowest
                              scenario).  Per-scenario iterates are the same
                              either way for EVERY_ITER and FIRST_ITER          */
  int32_t dense_block;     /* size of the sparse LU's dense trailing block, chosen
                              once per analysis: 0 (default) plans every bound in
                              16-row steps and keeps the one whose factorisation
                              and solve of scenario 0 are fastest on the device
                              (a timing: contexts may choose differently and then
                              round differently); -1 the schedule's cost model
                              alone (reproducible); k > 0 at most k rows       */
  int32_t retry_window;    /* an inexact Newton solve (factorisations reused, or
                              FP32 factors) whose |F|inf has not halved within
                              this many corrections, or that reaches itermax,
                              restarts from the step's initial iterate as
                              exact Newton with FP64 factors (the reference's
                              method, run to itermax) before its scenario
                              counts as failed: 0 (default) 8, -1 at itermax
                              only                                              */
  int32_t global_x;        /* 1: the batched triangular solves keep the solution in
                              global memory even where it fits in LDS (the form
                              they take by themselves above about 12k rows)    */
  double refactor_theta;   /* ADAPTIVE / ASYNC: refactorise when an iteration
                              reduces |F|inf by less than this factor (0: 30)   */
  double async_below;      /* ASYNC: a refactorisation asked for at |F|inf below
                              this runs on the second stream, above it in the
                              lock step (0: every one runs on the second stream) */
  int32_t predictor;       /* first Newton iterate of a step: 0 (default) z_k,
                              the reference's start (src/dynamics.jl:548-640);
                              1: the linear extrapolation 2 z_k - z_(k-1) under
                              the policies that reuse factorisations (ADAPTIVE,
                              ASYNC; not across an event, whose step starts
                              from z_k); 2: the quadratic one
                              3 z_k - 3 z_(k-1) + z_(k-2) once two steps have
                              been accepted since the start or the last event
                              (linear before).  Extrapolated starts save
                              corrections but may converge to another solution
                              of a step's equations than the reference's start
                              (one of 225 scenarios on the 2,000-bus workload).
                              A solve that fails restarts from z_k as exact
                              Newton either way                              */
  int32_t pivot_rule;      /* static pivots of gp_integrate's batched LU plans:
                              0 (default) large in the intact matrix and in the
                              matrix of every distinct set of tripped lines; 1
                              the intact matrix alone.  Either way, in the
                              free-running loop (lock_step 0), a scenario whose
                              factorisation meets an exactly zero pivot
                              (klu!'s SingularException, src/numerics.jl:144-167)
                              continues on a pivot sequence chosen for its own
                              matrix (gp_stats.repivots); lock step and the
                              adjoint report GP_ERR_SINGULAR_PIVOT instead    */
  int32_t solve_levels;    /* triangular solves of the batched LU plans whose
                              solution fits in LDS: 0 (default) partitioned,
                              up to 4 levels of the elimination per partition
                              step; 1 one level per step; k > 1 up to k levels
                              per partition.  A partition's rows are finished
                              from the values before it through composite
                              entries that each factorisation computes */
} gp_options;

typedef struct gp_stats {
  int64_t newton_iters;        /* total Newton iterations summed over scenarios   */
  int64_t scenario_steps;      /* sum over scenarios of integrated steps          */
  int64_t event_resolves;      /* dt = 0 re-solves                                 */
  int64_t refactorizations;    /* numeric factorizations of S or J (x scenarios)  */
  int32_t n_not_converged;     /* scenarios with at least one failed Newton solve */
  int32_t n_nan;               /* scenarios that produced a non-finite residual   */
  double max_final_residual;   /* max over scenarios/steps of the accepted |F|inf */
  double elapsed_ms;           /* device time of the integration (HIP events)     */
  double solve_ms;             /* device time inside the integrator's Newton loop */
  int64_t newton_rounds;       /* lock-step Newton iterations (launch rounds)      */
  int64_t factor_launches;     /* rounds that launched a numeric factorisation     */
  /* with gp_options.profile: device time, launches and scenarios served of
     the batched triangular solve and of the numeric factorisation, and their
     algorithmic HBM bytes per scenario (solve: b_f*nnz(L+U) + 32*rows for the
     factor values, rhs, solution and state update; factorisation:
     8*nnz(A) + b_f*nnz(L+U) for the matrix values read and the factors
     written; b_f = 4 or 8 bytes per factor value)                          */
  double solve_kernel_ms;
  int64_t solve_kernel_launches, solve_kernel_scenarios;
  double factor_kernel_ms;
  int64_t factor_kernel_launches, factor_kernel_scenarios;
  double solve_bytes_per_scenario, factor_bytes_per_scenario;
  /* Newton solves run again as exact Newton: under a refactorisation policy
     that reuses factorisations (FIRST_ITER, ADAPTIVE, ASYNC) a solve that
     fails restarts from the step's initial iterate with a factorisation at
     every iteration (the reference's method) before the scenario counts as
     failed                                                                 */
  int64_t exact_retries;
  /* scenarios that met an exactly zero pivot and continued on a pivot
     sequence analysed for their own matrix (the reference's fresh klu after
     a SingularException, src/numerics.jl:144-167): each continues from the
     start of the failed step, alone, on that plan                          */
  int64_t repivots;
} gp_stats;

typedef struct gp_ctx gp_ctx;

/* Build a context for M scenarios: copies tables, builds the Jacobian / Schur
 * patterns, the symbolic LU analyses (ordering, pivot order, level schedule)
 * and allocates device memory. */
int gp_create(const gp_system_desc *desc, int32_t M, const gp_options *opt,
              gp_ctx **out);
void gp_destroy(gp_ctx *ctx);

/* z0: [n] (shared) or [M*n] scenario-major rows when per_scenario != 0;
 * p: [n_par], shared by all scenarios. */
int gp_set_state(gp_ctx *ctx, const double *z0, int32_t per_scenario,
                 const double *p);

/* Per-scenario faults and horizons (nsteps[m] = round(tf_m/dt)).  Clears
 * the line trips of gp_set_line_trips.  At most 8 events in force at once
 * per scenario (GP_ERR_UNSUPPORTED otherwise). */
int gp_set_events(gp_ctx *ctx, const gp_event *ev, int32_t n_ev,
                  const int32_t *nsteps);

/* Per-scenario line trips and reclosures, beside the faults of the last
 * gp_set_events (call it after that).  Faults, tripped lines and dropped
 * loads share 8 event slots per scenario: a set of events that would have
 * more than 8 in force at once in a scenario is refused here
 * (GP_ERR_UNSUPPORTED), before any integration.  The LU plans choose pivots
 * large in the intact matrix and in the matrix of every distinct set of
 * lines out (each tripped line alone, and all of a scenario's at once).
 * The reference runs these events in its CPU integrator only
 * (src/dynamics.jl:502-655); here every method takes them, in the forward
 * sweep and in gp_adjoint. */
int gp_set_line_trips(gp_ctx *ctx, const gp_line_trip *trips, int32_t n_trips);

/* Per-scenario ZIP load disconnects (call after gp_set_events, which clears
 * them).  They share the 8 event slots of a scenario with its active faults
 * and tripped lines, hold a slot from their step on, and are refused here
 * when a scenario would have more than 8 events in force at once. */
int gp_set_load_drops(gp_ctx *ctx, const gp_load_drop *drops, int32_t n_drops);

/* Per-scenario device disconnects (ZIP loads and exciters, gp_device_drop;
 * call after gp_set_events, which clears them; replaces the drops of an
 * earlier gp_set_load_drops / gp_set_device_drops).  Same event slots and
 * limits as gp_set_load_drops.  Every method takes them, forward and
 * gp_adjoint (an offline exciter's parameters then have no pullback). */
int gp_set_device_drops(gp_ctx *ctx, const gp_device_drop *drops, int32_t n_drops);

/* Integrate all scenarios from the current state to their horizons.
 * tol: Newton tolerance on |F|inf per scenario; itermax per Newton solve. */
int gp_integrate(gp_ctx *ctx, int32_t method, double dt, double tol,
                 int32_t itermax, gp_stats *stats);

/* Recorded trajectory of scenario m: out[j*ld + k] for recorded channel j and
 * step k = 0..nsteps[m]; ld >= nsteps[m]+1. */
int gp_get_trajectory(gp_ctx *ctx, int32_t m, double *out, int64_t ld);

/* Outcome of the last gp_integrate per scenario: status[m] is GP_OK or the
 * gp_status that stopped scenario m (a Newton solve that did not converge,
 * a non-finite residual, a zero pivot); steps_done[m] is the number of time
 * steps completed before it stopped (nsteps[m] when it finished), so columns
 * 0..steps_done[m] of its trajectory are converged states.  The reference
 * ignores a failed Newton solve and continues (src/dynamics.jl:593-597,
 * src/batched.jl:459, ext:3044); this library stops the scenario and says so.
 * Either output may be NULL. */
int gp_get_scenario_status(gp_ctx *ctx, int32_t *status, int32_t *steps_done);

/* Final state of every scenario, out[m*n + j]. */
int gp_get_state(gp_ctx *ctx, double *out);

/* Residual F(z) of the backward-Euler system at z (scenario-major [M*n]) with
 * zold and step dt, with scenario m's faults active when fault_on[m] != 0.
 * Output f is [M*n]. */
int gp_eval_residual(gp_ctx *ctx, const double *z, const double *zold,
                     double dt, const uint8_t *fault_on, double *f);

/* Allocation-free residual.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * The same residual as gp_eval_residual, but every array is in device
 * memory in the library's layout -- row i of scenario m at i*ld + m, with
 * ld = gp_symbolic_info.ld_scen (M rounded up to 64) -- and the call only
 * launches kernels on `stream` (NULL: the context's stream) and returns: no
 * allocation, copy or synchronisation.  z, zold: [n*ld] inputs; f: [n*ld]
 * output (rows of scenarios m >= M untouched); fault_on: [ld] device flags
 * or NULL, with gp_eval_residual's meaning.  The library's device memory
 * for the call is a scratch array of GP_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race.  A context whose evaluation needs more
 * (ld + 12*MAXFAULT*ld + 16*n_clusters*ld bytes, 256-byte aligned parts) is
 * refused with GP_ERR_UNSUPPORTED.  Needs gp_set_state (the parameters). */
#define GP_NOALLOC_SCRATCH_BYTES (4u << 20)
int gp_eval_residual_noalloc(gp_ctx *ctx, const double *z, const double *zold, double dt,
                             const uint8_t *fault_on, double *f, void *stream);

/* Jacobian values of the backward-Euler system at z, in the library's CSR
 * pattern (gp_jacobian_pattern), out [M*nnz]. */
int gp_jacobian_pattern(gp_ctx *ctx, int32_t *nnz, int32_t *rowptr,
                        int32_t *colind);
int gp_eval_jacobian(gp_ctx *ctx, const double *z, double dt,
                     const uint8_t *fault_on, double *jval);

/* One Newton correction dz (z_new = z + dz) per scenario at (z, zold, dt)
 * with the given method, scenario m's faults active when fault_on[m] != 0.
 * Inputs/outputs are [M*n].  The context's integration state is replaced
 * by z.  Test hook for the Schur algebra: every method returns -J^{-1}F. */
int gp_newton_step(gp_ctx *ctx, int32_t method, const double *z, const double *zold,
                   double dt, const uint8_t *fault_on, double *dz);

/* ---- discrete adjoint ------------------------------------------------------
 * Trajectory objective: the reference's `functional` (src/dynamics.jl:330-337,
 * sum of squared GENROU speeds) summed over the steps with weight dt as in
 * test/test_adjoint_event_offset.jl:58-69, and over the scenarios:
 *   J = scale * sum_m sum_{k=1..nsteps[m]} dt * sum_r (z_k[rows[r]] - offset)^2
 * where z_k is the converged state of step k (before the step's event
 * re-solve).  rows == NULL: the GENROU speed deviations (src/utils.jl:46-59);
 * scale == 0: 1 (the reference's sum); 1/n_rows gives the mean. */
typedef struct gp_objective {
  int32_t n_rows;
  const int32_t *rows;
  double scale;
  double offset;
} gp_objective;

typedef struct gp_adjoint_stats {
  double elapsed_ms;        /* device time of the whole call                      */
  double forward_ms;        /* the checkpointed forward sweep                      */
  double backward_ms;       /* segment recomputation and the backward sweep       */
  int64_t scenario_steps;   /* forward scenario-steps                              */
  int64_t adjoint_steps;    /* backward scenario-steps (transposed solves)         */
  int32_t n_segments;       /* checkpoint segments                                 */
  int32_t n_excluded;       /* scenarios left out (their forward sweep failed)    */
  int64_t newton_iters;     /* Newton corrections, recomputation included         */
  int64_t reserved[4];
} gp_adjoint_stats;

/* Gradient of J w.r.t. the parameters p[params[j]] and the initial state by
 * the discrete adjoint of the backward-Euler sweep, the initial state held
 * fixed when p moves.  Replaces GradPower.adjoint(λ0, dprob, sys, traj, tvec;
 * functional=true) -> (λ, μ, λ_traj) of examples/adjoint.jl:34-51,106, a
 * function the reference snapshot calls but no longer contains.
 * The forward sweep (method, dt, tol, itermax as in gp_integrate) starts from
 * the state of the last gp_set_state with the events of gp_set_events, keeps
 * a checkpoint every `segment` steps (0: 30) and recomputes each segment
 * before its backward pass.  Backward step k solves
 *   (dF_k/dz_k)^T lam_k = dJ/dz_k + lam_{k+1} on the differential rows
 * with the Jacobian of forward step k (its fault state: the alignment of
 * test/test_adjoint_event_offset.jl:3-24) and accumulates
 * dJ/dp -= lam_k^T dF_k/dp.  Parameters with a pullback: GENROU H, D; SEXS K;
 * ESDC1A KA; TGOV1 R; IEESGO K1; IEEEST KS (others: GP_ERR_UNSUPPORTED).
 * Outputs (all but grad may be NULL): value[m] objective of scenario m,
 * grad[j] = dJ/dp[params[j]], grad_scen[m*n_params + j] scenario m's share,
 * dz0[m*n + i] = dJ/dz0 of scenario m.  A scenario whose forward sweep fails
 * is left out (zero outputs) and the call returns that failure's status
 * after filling the outputs.  Afterwards gp_get_trajectory and gp_get_state
 * give the forward sweep's trajectory and final state. */
int gp_adjoint(gp_ctx *ctx, int32_t method, double dt, double tol, int32_t itermax, int32_t segment,
               const gp_objective *obj, int32_t n_params, const int32_t *params, double *value,
               double *grad, double *grad_scen, double *dz0, gp_adjoint_stats *stats);

/* Sizes of the symbolic analyses (for the roofline byte counts). */
typedef struct gp_symbolic_info {
  int32_t n, nnz_j, nnz_lu_j, levels_j;
  int32_t n_red, nnz_s, nnz_lu_s, levels_s;
  int32_t n_clusters_nt, w_total_nt;
  int32_t nd_j, nd_s, nd_jt;   /* dense trailing block of each LU plan (autotuned) */
  int32_t res_j, res_s;        /* 1: the plan's factorisation keeps its multipliers in LDS
                                  (at the element size it was last shaped for)       */
  int32_t ld_scen;             /* row stride of the device arrays: M rounded up to 64 */
  int32_t solve_levels_j;      /* levels per partition of the J plan's solves (1: one level per step) */
  int32_t solve_levels_s;      /* the same for the S plan */
} gp_symbolic_info;
int gp_get_symbolic_info(gp_ctx *ctx, gp_symbolic_info *info);

/* Pivot analysis of the batched LU plans, on the host (no GPU): the column
 * order (minimum degree on A+A^T) and the static pivot row of every step,
 * chosen by threshold partial pivoting (KLU's rule, diagonal preferred
 * within `threshold`, the reference's klu(J), src/dynamics.jl:576-580) on
 * the CSC matrix (colptr, rowind, val) and, with n_extra > 0, large in every
 * extra representative matrix of the same pattern as well (the library
 * passes one per distinct set of tripped lines).  col_order[k] is the column
 * and row_of_pivot[k] the row of pivot k, 0-based.  Returns
 * GP_ERR_SINGULAR_PIVOT when the matrix has no nonzero candidate at a step. */
int gp_analyse_pivots(int32_t n, const int32_t *colptr, const int32_t *rowind, const double *val,
                      int32_t n_extra, const double *const *extra, double threshold, int32_t *col_order,
                      int32_t *row_of_pivot);

/* Measurement helper with no counterpart in the reference: a plain device
 * copy kernel dst[i] = src[i], i < n (device pointers), on the context's
 * stream; used to calibrate hardware byte counters against a known size. */
int gp_copy_kernel(gp_ctx *ctx, const double *src, double *dst, int64_t n);

const char *gp_last_error(void);
const char *gp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GRADPOWER_HIP_H */
