/*
 * bacon_gpu.h — C ABI of the MI355X-native batched IVP engine (libbacon_gpu.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an int status:
 * 0 = OK, 1..12 = the reference's IVPError variants in declaration order
 * (src/ivp.rs:50-76), negative = misuse of this C interface or a HIP failure.
 *
 * Two families:
 *   bg_*           single-instance builder + iterator mirroring the IVPSolver
 *                  builder (src/ivp.rs:134-190) and IVPIterator
 *                  (src/ivp.rs:192-239).  The whole path is produced on the
 *                  GPU by one lane; bg_next walks it on the host.
 *   bg_ensemble_*  the batched builder: the same setters taking per-instance
 *                  arrays, solved by persistent gfx950 kernels.
 * plus device-pointer entry points (bg_*_device) for callers whose inputs
 * already live in HBM.
 *
 * Right-hand sides are registered device functions selected by id (the
 * Derivative closure of src/ivp.rs:34-48 cannot cross a C ABI into GPU code);
 * per-instance parameters replace the reference's UserData.
 */
#ifndef BACON_GPU_H
#define BACON_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BG_OK 0
#define BG_MISSING_PARAMETERS 1          /* IVPError::MissingParameters       */
#define BG_USER_ERROR 2                  /* IVPError::UserError               */
#define BG_TOLERANCE_OOB 3               /* IVPError::ToleranceOOB            */
#define BG_TIME_DELTA_OOB 4              /* IVPError::TimeDeltaOOB            */
#define BG_TIME_END_OOB 5                /* IVPError::TimeEndOOB              */
#define BG_TIME_START_OOB 6              /* IVPError::TimeStartOOB            */
#define BG_FROM_PRIMITIVE_FAILURE 7      /* IVPError::FromPrimitiveFailure    */
#define BG_MINIMUM_TIME_DELTA_EXCEEDED 8 /* IVPError::MinimumTimeDeltaExceeded */
#define BG_MAXIMUM_ITERATIONS_EXCEEDED 9 /* IVPError::MaximumIterationsExceeded*/
#define BG_SINGULAR_MATRIX 10            /* IVPError::SingularMatrix          */
#define BG_DYNAMIC_ON_STATIC 11          /* IVPError::DynamicOnStatic         */
#define BG_STATIC_ON_DYNAMIC 12          /* IVPError::StaticOnDynamic         */
#define BG_DONE 100                      /* iterator exhausted (IVPStatus::Done) */
#define BG_ERR_INVALID_ARGUMENT (-1)     /* null pointer, size mismatch, ...  */
#define BG_ERR_HIP (-2)                  /* a HIP runtime call failed         */
#define BG_ERR_UNSUPPORTED (-3)          /* method/RHS/dimension not compiled */
#define BG_ERR_STATE (-4)                /* call out of order (e.g. next before solve) */

/* ---- methods (coefficient plugins rk.rs:20-45 / bdf.rs:20-41) ---------- */
#define BG_RK45 0 /* RungeKutta45, rk.rs:561 */
#define BG_RK23 1 /* RungeKutta23, rk.rs:656 */
#define BG_BDF6 2 /* BDF6, bdf.rs:706        */
#define BG_BDF2 3 /* BDF2, bdf.rs:762        */
#define BG_ADAMS5 4 /* Adams5, adams.rs:633 */
#define BG_ADAMS3 5 /* Adams3, adams.rs:693 */

/* ---- semantics: the reference's quirk register as data ------------------
 * BG_AS_BUILT : the shipped behaviour (RK stage matrix read transposed,
 *               FD Jacobian column (G+ + G-)/2dt, lower residual with the
 *               higher history weights, f at t_n, bootstrap time restore
 *               t -= dt - 7).
 * BG_INTENDED : the documented methods (Fehlberg rows as rows, central
 *               difference, own BDF5 weights, f at t_n+1, t -= 7 dt).
 * Every other bacon semantic (safety factor 1.0, clamps, dt_min failure,
 * RK4 bootstrap, Broyden corrector) is shared. */
#define BG_AS_BUILT 0
#define BG_INTENDED 1

/* ---- dimension kinds (Dimension trait, lib.rs:52-76) -------------------- */
#define BG_DIM_STATIC 0 /* nalgebra Const<D> */
#define BG_DIM_DYN 1    /* nalgebra Dyn      */

/* ---- registered right-hand sides (device functions) -------------------- */
#define BG_RHS_EXP 0            /* y' = y (bdf.rs:769-771)                   */
#define BG_RHS_NEG 1            /* y' = -y (bdf.rs:781-783)                  */
#define BG_RHS_QUADRATIC 2      /* y' = -2t (rk.rs:664-666)                  */
#define BG_RHS_COS 3            /* y' = cos t (rk.rs:668-670)                */
#define BG_RHS_LORENZ 4         /* params: sigma, rho, beta; dim 3           */
#define BG_RHS_LINEAR 5         /* y' = A y + b, shared [A row-major; b];
                                   dim 32, 64 or 128; RK45                   */
#define BG_RHS_BRUSSELATOR 6    /* params: A, B, alpha; dim 2*cells          */
#define BG_RHS_LOTKA_VOLTERRA 7 /* params: alpha, beta, gamma, delta; dim 2  */
#define BG_RHS_OSCILLATOR 8     /* y0' = y1, y1' = -y0 (ivp.rs cos_deriv)    */
#define BG_RHS_FAIL_AFTER 9     /* y' = y; UserError once t >= params[0]     */
#define BG_RHS_ROBERTSON 10     /* stiff kinetics; params: k1, k2, k3; dim 3 */
/* Complex state (the reference's N: ComplexField, rk.rs:53, adams.rs:44): a
 * system of m complex components is 2m reals, each component stored as its
 * (re, im) pair; RK and Adams only (BDF's Broyden update multiplies by the
 * plain transpose, bdf.rs:461-466, so its complex run is not the real-pair
 * run).  Two-level system c' = (-i H - gamma) c, H = [[hd, ho], [ho, -hd]]:
 * params hd, ho, gamma; dim 4 (two complex components). */
#define BG_RHS_TWO_LEVEL 11

/* ======================================================================== */
/* Single-instance builder + iterator                                       */
/* ======================================================================== */
typedef struct bg_solver bg_solver;

/* RungeKutta45::new() / BDF6::new() (rk.rs:136-148, bdf.rs:144-156) for a
 * solver whose dimension type is `dim_kind`; `dim` is the static size.
 * A BG_DIM_DYN kind fails with BG_STATIC_ON_DYNAMIC like Dyn::dim(). */
int bg_new(int method, int dim_kind, uint32_t dim, bg_solver **out);
/* new_dyn(size) (rk.rs:150-162): a BG_DIM_STATIC kind fails with
 * BG_DYNAMIC_ON_STATIC like Const::dim_dyn(). */
int bg_new_dyn(int method, int dim_kind, uint32_t size, bg_solver **out);
/* IVPSolver::dim (ivp.rs:166) */
int bg_dim(const bg_solver *s, uint32_t *dim);
/* setters: rk.rs:168-247 / bdf.rs:176-255, same validation and order */
int bg_with_tolerance(bg_solver *s, double tol);
int bg_with_maximum_dt(bg_solver *s, double max);
int bg_with_minimum_dt(bg_solver *s, double min);
int bg_with_initial_time(bg_solver *s, double initial);
int bg_with_ending_time(bg_solver *s, double ending);
/* with_initial_conditions_slice (ivp.rs:177-180); n must equal the dimension */
int bg_with_initial_conditions_slice(bg_solver *s, const double *y0, size_t n);
/* with_derivative (rk.rs:244-247): a registered device RHS + its parameters */
int bg_with_derivative(bg_solver *s, int rhs, const double *params, size_t n_params);
/* quirk table selection (default BG_INTENDED) */
int bg_with_semantics(bg_solver *s, int semantics);
/* README.md:33-38 spellings of the same setters */
int bg_with_dt_min(bg_solver *s, double min);
int bg_with_dt_max(bg_solver *s, double max);
int bg_with_start(bg_solver *s, double initial);
int bg_with_end(bg_solver *s, double ending);
/* solve(data) (rk.rs:249-343, bdf.rs:257-331, adams.rs:249-333): checks
 * MissingParameters, then produces the path on the GPU (one lane; for Adams
 * the iterator's yields, bootstrap replays of adams.rs:415-428 included). */
int bg_solve(bg_solver *s);
/* IVPIterator::next (ivp.rs:220-238): BG_OK with (t, y[dim]) filled,
 * BG_DONE at the end, or the failure code once (then BG_DONE: fused). */
int bg_next(bg_solver *s, double *t, double *y);
/* IVPStepper::time (ivp.rs:124): time of the last yielded value */
int bg_time(const bg_solver *s, double *t);
void bg_destroy(bg_solver *s);

/* ======================================================================== */
/* Batched builder                                                          */
/* ======================================================================== */
typedef struct bg_ensemble bg_ensemble;

/* counters per instance */
typedef struct {
    uint32_t n_accepted, n_rejected;
} bg_counts;

int bg_ensemble_new(int method, uint32_t dim, uint64_t n, bg_ensemble **out);
/* per-instance setters: `count` is 1 (broadcast) or n.  The validation of
 * each element is the single-instance setter's. */
int bg_ensemble_with_tolerance(bg_ensemble *e, const double *tol, uint64_t count);
int bg_ensemble_with_maximum_dt(bg_ensemble *e, const double *v, uint64_t count);
int bg_ensemble_with_minimum_dt(bg_ensemble *e, const double *v, uint64_t count);
int bg_ensemble_with_initial_time(bg_ensemble *e, const double *v, uint64_t count);
int bg_ensemble_with_ending_time(bg_ensemble *e, const double *v, uint64_t count);
/* y0 structure-of-arrays: y0[d*n + i]; count = dim*n */
int bg_ensemble_with_initial_conditions(bg_ensemble *e, const double *y0, uint64_t count);
/* params structure-of-arrays: params[k*n + i], k < n_params, i < n, so the
 * array holds n_params * n doubles; shared: RHS-wide data of n_shared
 * doubles */
int bg_ensemble_with_derivative(bg_ensemble *e, int rhs, const double *params,
                                uint32_t n_params, const double *shared, uint64_t n_shared);
int bg_ensemble_with_semantics(bg_ensemble *e, int semantics);
/* cap on step attempts per instance (guards the reference's endless Redo on
 * a NaN error); 0 = default */
int bg_ensemble_with_max_attempts(bg_ensemble *e, uint64_t max_attempts);
/* BDF: arithmetic of the Jacobian inverse, 64 (default) or 32 bits */
int bg_ensemble_with_lu_precision(bg_ensemble *e, int bits);
/* Solve on the GPU (synchronous). */
int bg_ensemble_solve(bg_ensemble *e);
/* Plain path: the same device step code launched once per instance,
 * host-synchronous, one instance after another. */
int bg_ensemble_solve_plain(bg_ensemble *e);
/* solve_ivp (README.md:45-47) per instance: Adams5, then RK45 for the
 * instances that did not finish OK, then BDF6 for those still failing
 * (methods that do not support the right-hand side are skipped).  The
 * ensemble's own method is not used; its semantics are.  method_used[n]
 * (may be NULL) receives the method of each instance's final result, -1 if
 * none applied.  Results as from bg_ensemble_solve. */
int bg_ensemble_solve_ivp(bg_ensemble *e, int32_t *method_used);
/* Copy results to host arrays (any may be NULL); y_out is [dim][n]. */
int bg_ensemble_results(bg_ensemble *e, double *t_out, double *y_out, int32_t *status,
                        bg_counts *counts);
/* BDF work counters per instance: (Broyden iterations, Jacobian inverses) */
int bg_ensemble_work(bg_ensemble *e, bg_counts *work);
void bg_ensemble_destroy(bg_ensemble *e);

/* ======================================================================== */
/* Device-pointer entry points (inputs already resident in HBM)             */
/* ======================================================================== */

/* Launch shape chosen by the autotuner (zeros = library defaults). */
typedef struct {
    uint32_t block;           /* threads per block (multiple of 64)          */
    uint32_t blocks_per_cu;   /* persistent blocks per CU (0 = occupancy)    */
    uint32_t refill_min;      /* refill a wave once this many lanes are idle */
    uint32_t prio_cutoff;     /* waves holding one of the first prio_cutoff
                                 scheduled instances run at raised priority */
} bg_launch;

typedef struct {
    uint64_t n;
    uint32_t dim, n_params;
    /* inputs (device pointers; per-instance arrays may be NULL = scalar) */
    const double *y0;      /* [dim][n]        */
    const double *params;  /* [n_params][n]   */
    const double *shared;  /* RHS-wide block  */
    const double *t0, *t_end, *tol, *dt_min, *dt_max; /* [n] or NULL */
    double t0_s, t_end_s, tol_s, dt_min_s, dt_max_s;  /* scalar fallbacks */
    const int32_t *order;  /* schedule permutation [n] or NULL */
    uint64_t max_attempts; /* 0 = default */
    /* outputs (device pointers; any may be NULL) */
    double *t_out;         /* [n]             */
    double *y_out;         /* [dim][n]        */
    int32_t *status;       /* [n]             */
    bg_counts *counts;     /* [n]             */
    bg_counts *work;       /* [n] BDF only: (Broyden iterations, Jacobian inverses) */
} bg_ensemble_desc;

/* Batched RK (RK45 / RK23) in one persistent launch on `stream`
 * (hipStream_t, NULL = default stream).  Asynchronous. */
int bg_rk_solve_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                       const bg_launch *launch, void *stream);
/* The same launch without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * launch's workspace (work counter and descriptor copy, 1 KiB at most: it
 * does not grow with d->n) is a static device array of the library, one per
 * GPU, reused by every call on that GPU: calls that overlap on one GPU race.
 * Same arguments, results and errors as bg_rk_solve_device.  Asynchronous. */
int bg_rk_solve_device_noalloc(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                               const bg_launch *launch, void *stream);
/* The plain per-instance path over the same descriptor (synchronous). */
int bg_rk_solve_plain_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                             void *stream);
/* Path output for an ensemble: every accepted (t, y) that the reference's
 * IVPIterator yields (rk.rs:419, ivp.rs:220-238), per instance, into a ring
 * buffer of `capacity` entries (entry k of an instance at k % capacity;
 * path_len[i] counts all entries, so > capacity means the ring wrapped).
 * Optional shared stop times: the end clamp (rk.rs:366-368) aims at the next
 * stop and an accepted clamped step lands exactly on it. */
#define BG_PATH_STOPS_ONLY 1u   /* record only landings on stops and the end */
/* Instance-major buffers: path_t[n][capacity], path_y[n][capacity][dim].
 * Each instance's records are contiguous, so the stores of consecutive
 * steps fill the same cache lines; the default [capacity][n] layout
 * coalesces only while the lanes of a wave record in step, and instances
 * desynchronise within a few steps (measured on 262,144 Lorenz-63
 * instances with every step recorded: 41 ms default, see the bench). */
#define BG_PATH_INSTANCE_MAJOR 2u
typedef struct {
    uint32_t capacity;     /* entries per instance (0: count only)             */
    uint32_t every;        /* record every k-th accepted step (0 or 1: each)   */
    uint32_t n_stops;
    uint32_t flags;        /* BG_PATH_STOPS_ONLY | BG_PATH_INSTANCE_MAJOR      */
    const double *stops;   /* [n_stops] increasing times shared by all, or NULL */
    double *path_t;        /* [capacity][n] (instance-major: [n][capacity])    */
    double *path_y;        /* [capacity][dim][n] ([n][capacity][dim])          */
    uint32_t *path_len;    /* [n] entries recorded, or NULL                    */
} bg_path_desc;
/* Batched RK with path output, one persistent launch, asynchronous. */
int bg_rk_solve_path_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                            const bg_path_desc *path, void *stream);

/* ------------------------------------------------------------------------ */
/* Autotuner (no counterpart in the reference: its solver runs one instance) */
/* ------------------------------------------------------------------------ */
#define BG_TUNE_MAX_PARAMS 8
/* A tuned plan for one RK workload.
 * Schedule: instances are handed to lanes longest-first by the predicted cost
 *   log cost = coef[0] + coef[1] log(t_end - t0) + coef[2] log(tol)
 *            + sum_k coef[3 + k] params[k]
 * whose coefficients are a least-squares fit to the step attempts of a solve
 * of the workload (on a subsample of up to 16,384 instances).
 * Launch: the fastest of a set of (block size, resident blocks per CU)
 * shapes, each timed on the whole problem with that schedule. */
typedef struct {
    bg_launch launch;
    uint32_t n_params;   /* parameter terms in the model                     */
    uint32_t pilot_n;    /* instances the model was fitted on                */
    double coef[3 + BG_TUNE_MAX_PARAMS];
    double pilot_ms;     /* the solve in input order, default shape          */
    double default_ms;   /* default shape with the fitted schedule           */
    double best_ms;      /* the chosen shape with the fitted schedule        */
    uint32_t n_trials;
    uint32_t pad;
} bg_tuning;

/* Tune `d` (d->order must be NULL).  Runs 1 + n_trials full solves on
 * `stream`, each of which writes valid results into d's outputs; returns
 * when they are done. */
int bg_rk_autotune_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                          bg_tuning *out, void *stream);
/* The plan's schedule permutation of d's instances into order[n] (device). */
int bg_rk_schedule_device(const bg_tuning *tuning, const bg_ensemble_desc *d, int32_t *order,
                          void *stream);
/* Longest-first permutation order[n] (device) from an earlier solve's
 * per-instance work counts[n] (device; accepted + rejected attempts), e.g.
 * the previous iteration of an optimisation loop over the same instances.
 * Stable: equal counts keep index order.  For d->order of the RK ensemble or
 * of the adjoint's forward pass.  No counterpart in the reference, like the
 * autotuner. */
int bg_schedule_by_counts_device(const bg_counts *counts, uint64_t n, int32_t *order,
                                 void *stream);
/* Schedule + solve with the plan's launch shape, asynchronous on `stream`.
 * If d->order is set it is used as given. */
int bg_rk_solve_tuned_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                             const bg_tuning *tuning, void *stream);

/* ------------------------------------------------------------------------ */
/* HIP graph capture (no counterpart in the reference)                      */
/* ------------------------------------------------------------------------ */
/* One batched solve captured once and replayed: the setup kernel (counter
 * reset, descriptor copy) and the persistent solve kernel, whose BDF Newton
 * iterations run inside it.  The launch workspace is allocated at capture and
 * owned by the graph.  A replay reads the descriptor's device buffers as they
 * hold at replay time.  Replays of one graph must not overlap each other. */
typedef struct bg_graph bg_graph;
int bg_graph_capture_rk(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                        const bg_launch *launch, bg_graph **out);
int bg_graph_capture_bdf(int method, int rhs, int semantics, int lu_bits,
                         const bg_ensemble_desc *d, bg_graph **out);
/* Replay on `stream` (hipStream_t, NULL = default), asynchronous. */
int bg_graph_launch(bg_graph *g, void *stream);
void bg_graph_destroy(bg_graph *g);

/* Batched BDF (BDF6 / BDF2) in one persistent launch on `stream`
 * (asynchronous).  lu_bits = 64: Jacobian inverse in FP64; 32: inverse
 * factored in FP32, every residual and the Broyden iteration in FP64. */
int bg_bdf_solve_device(int method, int rhs, int semantics, int lu_bits,
                        const bg_ensemble_desc *d, void *stream);

/* The autotuner for a BDF workload (no counterpart in the reference): the
 * schedule model of bg_tuning fitted to the corrector's work per instance
 * (16 x Jacobian inverses + Broyden iterations; d->work is filled if set),
 * then the cap on resident waves per CU (tuning->launch.blocks_per_cu, 0 =
 * occupancy; launch.block is 64) timed on the whole problem with that
 * schedule.  d->order must be NULL.  Runs 1 + n_trials full solves, each
 * writing valid results into d's outputs; returns when they are done. */
int bg_bdf_autotune_device(int method, int rhs, int semantics, int lu_bits,
                           const bg_ensemble_desc *d, bg_tuning *out, void *stream);
/* Schedule + solve with a BDF plan, asynchronous (d->order, if set, is used
 * as given). */
int bg_bdf_solve_tuned_device(int method, int rhs, int semantics, int lu_bits,
                              const bg_ensemble_desc *d, const bg_tuning *tuning, void *stream);

/* The BDF corrector's Jacobian inverse (bdf.rs:429-444) on `count` caller
 * matrices a[count][dim][dim] (row-major, device memory), dim in {1, 2, 4,
 * 8, 16, 32}: partial-pivot Gauss-Jordan, then full-pivot Gauss-Jordan,
 * then Householder QR, each tried when the one before meets an exact zero
 * pivot (nalgebra's lu / full_piv_lu / qr try_inverse order).  `first`
 * starts the cascade at member 0, 1 or 2 (0: the whole cascade, as the
 * corrector runs it).  inv gets the inverses (same layout) and
 * branch[count] the member that answered: 0, 1, 2, or -1 for
 * SingularMatrix.  No counterpart of its own in the reference: the code the
 * BDF kernels run inside the Broyden corrector, exposed for batched use and
 * checking.  Asynchronous on `stream`. */
int bg_bdf_inverse_device(uint32_t dim, uint64_t count, const double *a, double *inv,
                          int32_t *branch, int first, void *stream);

/* Batched Adams-Bashforth-Moulton (Adams5 / Adams3) in one persistent
 * launch on `stream` (asynchronous).  Semantics: BG_AS_BUILT keeps the
 * reference's derivative history, which misses the first Adams step after
 * each RK4 bootstrap (adams.rs:434-441); BG_INTENDED adds it. */
int bg_adams_solve_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                          void *stream);

/* Discrete adjoint of the accepted RK step sequence (RK45, intended tableau)
 * for the loss L = sum_k |y(t_k) - target_k|^2 over observation times t_k
 * shared by all instances (increasing, in (t0, t_end]; observation times
 * before t_end become stops: the step clamp of rk.rs:366-368 aims at the
 * next one and the state lands exactly on it).  The forward solve records a
 * checkpoint (t, dt, next stop, y) every `checkpoint_every` accepted steps
 * and the size of every accepted step; the backward pass re-runs each
 * segment's accepted steps from its checkpoint with those sizes
 * (bit-identical to the forward, without its rejected attempts) and applies
 * the reverse stage-by-stage VJP of the frozen step map.  Gradients are with respect to the per-instance parameters and
 * y0.  No counterpart in the reference (it has no sensitivity API). */
typedef struct {
    bg_ensemble_desc base;      /* inputs; status, counts, t_out, y_out (final) */
    uint32_t n_obs;
    uint32_t checkpoint_every;  /* K (0 = 16)                                   */
    uint32_t max_checkpoints;   /* per instance (0 = 128)                       */
    uint32_t blocks_per_cu;     /* upper bound on resident 256-thread blocks
                                   per CU for both passes (0: forward 2,
                                   backward 1; the backward's LDS tape allows
                                   one block per CU at K = 16); base.order,
                                   if set, is the forward pass's schedule       */
    const double *obs_t;        /* [n_obs]                                      */
    const double *targets;      /* [n_obs][dim][n]                              */
    double *loss;               /* [n]                                          */
    double *grad_params;        /* [n_params][n]                                */
    double *grad_y0;            /* [dim][n]                                     */
    double *obs_y;              /* [n_obs][dim][n] states at the observation
                                   times, or NULL                              */
} bg_adjoint_desc;
/* Loss and its gradient (vector-Jacobian product with a unit cotangent per
 * instance) on `stream`, asynchronous. */
int bg_rk_vjp_device(int method, int rhs, const bg_adjoint_desc *d, void *stream);

/* A tuned plan for the adjoint (no counterpart in the reference): the
 * checkpoint interval K, the resident blocks per CU and, unless the caller
 * gave the forward pass's order, whether it runs longest first by a cost
 * model fitted to the forward's attempts (the RK tuner's model: log
 * attempts linear in log(t_end - t0), log(tol) and the parameters).
 * Gradients depend on none of these (each instance's operations are the
 * same in any order, and each segment is re-run from its checkpoint with
 * the forward's operations), so the plan only changes the time. */
typedef struct {
    uint32_t checkpoint_every;  /* K of the fastest trial                       */
    uint32_t max_checkpoints;   /* scaled so that K x max_checkpoints covers the
                                   caller's accepted steps per instance        */
    uint32_t blocks_per_cu;     /* 0 = the launcher's defaults                 */
    uint32_t n_trials;
    double default_ms;          /* the caller's K and blocks_per_cu            */
    double best_ms;             /* the fastest trial, its schedule included    */
    uint32_t schedule;          /* 1: the forward pass runs in the order that
                                   bg_rk_schedule_device(&forward, &d->base, …)
                                   gives                                       */
    uint32_t pad;
    bg_tuning forward;          /* the fitted model (coef, n_params, pilot_n);
                                   its launch, pilot_ms and default_ms are 0   */
} bg_adjoint_tuning;
/* Times the gradient solve of `d` for K in {d's K, 16, 32, 48} and
 * blocks_per_cu in {d's, 0, 1, 2} (distinct pairs), in d's order and, when
 * d->base.order is NULL, also in the fitted longest-first order (the
 * schedule's cost kernel and sort inside the timed region); each trial
 * writes valid results into d's outputs.  Returns when they are done. */
int bg_adjoint_autotune_device(int method, int rhs, const bg_adjoint_desc *d,
                               bg_adjoint_tuning *out, void *stream);
/* bg_rk_vjp_device with the plan applied: its K, checkpoint capacity and
 * blocks per CU, and (plan->schedule, d->base.order NULL) the forward order
 * computed on the device from the fitted model, in a workspace of its own. */
int bg_adjoint_solve_tuned_device(int method, int rhs, const bg_adjoint_desc *d,
                                  const bg_adjoint_tuning *plan, void *stream);

/* Library / device information */
int bg_version(void);
int bg_device_info(int *n_cu, int *arch_major, int *arch_minor);

#ifdef __cplusplus
}
#endif
#endif
