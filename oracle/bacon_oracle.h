/*
 * bacon_oracle.h — CPU restatement of bacon-sci 0.16.2's adaptive IVP step
 * loop (RK45 / RK23 stage loop, embedded error estimate and step controller;
 * BDF6 / BDF2 RK4 bootstrap, residuals, finite-difference Jacobian, Broyden
 * corrector and controller).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in bacon_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only to compare against or to time the CPU
 * path.  Nothing in the product links, loads or calls it.
 *
 * Two arithmetic profiles share every code path:
 *   OR_PROFILE_BACON  : the reference's own arithmetic.  Rust does not fuse
 *                       multiply-adds, so every a*b+c is two roundings (the
 *                       file is compiled with -ffp-contract=off), and the
 *                       controller exponent is pow(x, 1/4) as in
 *                       src/ivp/rk.rs:400-401.
 *   OR_PROFILE_DEVICE : the same algorithm with the operation order of the
 *                       gfx950 kernels: explicit fma() at each multiply-add,
 *                       zero tableau entries skipped, and the controller on
 *                       squared norms (accept ss <= (tol dt)^2; step factor
 *                       r^(1/8) with r = (tol dt)^2 / ss from or_root8's
 *                       frexp / degree-5 polynomial / ldexp, relative error
 *                       below 1.11e-6, bounded by the tangent at r = 1).
 *                       Every operation is an IEEE correctly rounded
 *                       primitive or an exact bit operation on both CPU and
 *                       GPU, so the single-lane GPU RK kernels must match it
 *                       bit for bit.
 *
 * Pinning (tests/test_oracle.py): the reference's own ivp tests, order
 * conditions, closed forms, scipy golden vectors, LAPACK inverses.
 * Parity unpinned: the exact rounding of nalgebra 0.32's multi-dimensional
 * norm / LU / full-pivot LU / QR (restated from its published algorithms; the
 * crate is not vendored and cannot be built here) and as-built step counts for
 * y-dependent right-hand sides.
 */
#ifndef BACON_ORACLE_H
#define BACON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0 = OK, then IVPError's variants in declaration order
 * (src/ivp.rs:50-76) as 1..12. */
enum {
    OR_OK = 0,
    OR_MISSING_PARAMETERS = 1,
    OR_USER_ERROR = 2,
    OR_TOLERANCE_OOB = 3,
    OR_TIME_DELTA_OOB = 4,
    OR_TIME_END_OOB = 5,
    OR_TIME_START_OOB = 6,
    OR_FROM_PRIMITIVE_FAILURE = 7,
    OR_MINIMUM_TIME_DELTA_EXCEEDED = 8,
    OR_MAXIMUM_ITERATIONS_EXCEEDED = 9,
    OR_SINGULAR_MATRIX = 10,
    OR_DYNAMIC_ON_STATIC = 11,
    OR_STATIC_ON_DYNAMIC = 12,
    /* IVPStatus::Done / Redo (src/ivp.rs:20-28); never returned by or_solve */
    OR_STEP_DONE = 100,
    OR_STEP_REDO = 101,
    OR_STEP_YIELD = 102,
    /* misuse of this C interface (null pointer, bad dimension, ...) */
    OR_INVALID_ARGUMENT = -1
};

enum { OR_RK45 = 0, OR_RK23 = 1, OR_BDF6 = 2, OR_BDF2 = 3, OR_ADAMS5 = 4, OR_ADAMS3 = 5 };
enum { OR_AS_BUILT = 0, OR_INTENDED = 1 };
enum { OR_PROFILE_BACON = 0, OR_PROFILE_DEVICE = 1 };

/* Right-hand sides known to the oracle (own implementations, independent of
 * the device functors in bacon_amd/csrc/rhs.h). */
enum {
    OR_RHS_EXP = 0,          /* y' = y            (bdf.rs:769-771)            */
    OR_RHS_NEG = 1,          /* y' = -y           (bdf.rs:781-783)            */
    OR_RHS_QUADRATIC = 2,    /* y' = -2t          (rk.rs:664-666)             */
    OR_RHS_COS = 3,          /* y' = cos t        (rk.rs:668-670, bdf.rs:777) */
    OR_RHS_LORENZ = 4,       /* params sigma, rho, beta                       */
    OR_RHS_LINEAR = 5,       /* y' = A y + b, shared = [A row-major, b]       */
    OR_RHS_BRUSSELATOR = 6,  /* params A, B, alpha; dim = 2*cells, periodic   */
    OR_RHS_LOTKA_VOLTERRA = 7, /* params alpha, beta, gamma, delta            */
    OR_RHS_OSCILLATOR = 8,   /* y0' = y1, y1' = -y0 (ivp.rs cos_deriv)        */
    OR_RHS_FAIL_AFTER = 9,   /* y' = y, UserError once t >= params[0]         */
    OR_RHS_LV_SENS = 10,     /* 14-dim forward-sensitivity system of LV       */
    OR_RHS_ROBERTSON = 11,   /* stiff kinetics, params k1, k2, k3; dim 3      */
    OR_RHS_TWO_LEVEL = 12,   /* complex N: y' = (-i H - gamma) y, H = [[hd, ho],
                                [ho, -hd]], params hd, ho, gamma; two complex
                                components stored (re, im) pairs: dim 4       */
    OR_RHS_COUNT = 13
};

#define OR_MAX_DIM 128

typedef struct {
    int32_t method;      /* OR_RK45 ...                                    */
    int32_t semantics;   /* OR_AS_BUILT / OR_INTENDED (quirk table)        */
    int32_t profile;     /* OR_PROFILE_BACON / OR_PROFILE_DEVICE           */
    int32_t rhs;         /* OR_RHS_*                                       */
    int32_t dim;
    int32_t n_stops;     /* observation stops strictly inside (t0, t_end) */
    double t0, t_end, tol, dt_min, dt_max;
    int64_t max_attempts;   /* 0 -> default cap                            */
    const double *y0;       /* dim                                         */
    const double *params;   /* per-instance parameters                     */
    const double *shared;   /* shared data block (linear RHS: A, b)        */
    const double *stops;    /* sorted observation times                    */
} or_problem;

typedef struct {
    double t, dt;
    int32_t status;
    int32_t pad;
    int64_t n_accepted, n_rejected, n_rhs, n_newton, n_lu, n_yield;
} or_result;

/* Path recorder: every yielded (t, y) while it fits `cap` entries; `len`
 * receives the full number of yields.  For RK methods dt[k] is the step
 * size that produced entry k.  Any pointer may be NULL. */
typedef struct {
    int64_t cap, len;
    double *t, *y, *dt;
} or_path;

/* Solve one IVP.  y_out receives the solver state when the iterator ends
 * (Done) or fails.  stop_y (n_stops*dim, may be NULL) receives the state at
 * each observation stop. */
int32_t or_solve(const or_problem *p, double *y_out, or_result *r,
                 or_path *path, double *stop_y);

/* Ensemble driver.  Per-instance arrays are structure-of-arrays:
 * y0[d*n + i], params[k*n + i]; t_end / tol may be NULL (use the template).
 * Runs with OpenMP over `threads` threads (<=0: all). */
int32_t or_solve_ensemble(const or_problem *tmpl, int64_t n, int32_t n_params,
                          const double *y0, const double *params,
                          const double *t_end, const double *tol,
                          double *t_out, double *y_out, int32_t *status,
                          int64_t *n_accepted, int64_t *n_rejected,
                          int32_t threads);

/* Evaluate one RHS (for tests of the device functors). */
/* Norm of an error vector as the given profile's step computes it: BACON is
 * nalgebra's norm() (dotc, eight accumulators from 8 rows), over n reals or,
 * with complex_pairs, over n/2 complex components stored as (re, im) pairs
 * (each component contributes re^2 + im^2 as one term); DEVICE is the
 * sequential fused sum of squares over the n reals. */
double or_norm(int32_t n, const double *x, int32_t profile, int32_t complex_pairs);
/* 1 for a right-hand side with complex state (OR_RHS_TWO_LEVEL). */
int32_t or_rhs_is_complex(int32_t rhs);

int32_t or_rhs_eval(int32_t rhs, int32_t profile, int32_t dim, double t,
                    const double *y, const double *params,
                    const double *shared, double *out);

/* Builder restatement (IVPSolver setters, rk.rs:136-247, bdf.rs:144-255). */
typedef struct {
    int32_t method, is_dyn, dim, built;
    int32_t have_tol, have_dt_max, have_dt_min, have_t0, have_t_end, have_y0,
        have_deriv;
    double tol, dt_max, dt_min, t0, t_end;
} or_builder;
int32_t or_builder_new(or_builder *b, int32_t method, int32_t static_dim);
int32_t or_builder_new_dyn(or_builder *b, int32_t method, int32_t size);
int32_t or_builder_with_tolerance(or_builder *b, double tol);
int32_t or_builder_with_maximum_dt(or_builder *b, double v);
int32_t or_builder_with_minimum_dt(or_builder *b, double v);
int32_t or_builder_with_initial_time(or_builder *b, double v);
int32_t or_builder_with_ending_time(or_builder *b, double v);
int32_t or_builder_with_initial_conditions(or_builder *b, int32_t n);
int32_t or_builder_with_derivative(or_builder *b);
int32_t or_builder_solve(const or_builder *b);

/* nalgebra 0.32 inverse cascade of bdf.rs:429-444: partial-pivot LU, then
 * full-pivot LU, then Householder QR.  Returns OR_OK, or OR_SINGULAR_MATRIX.
 * a and inv are column-major n*n.  *which receives 0/1/2 for the method used. */
int32_t or_inverse(int32_t n, const double *a, double *inv, int32_t *which);
int32_t or_lu_inverse(int32_t n, const double *a, double *inv);
/* one member of the cascade: branch 0 = LU, 1 = full-pivot LU, 2 = QR */
int32_t or_inverse_branch(int32_t n, const double *a, double *inv, int32_t branch);

/* Tableau exposure for order-condition tests: semantics picks the stage
 * matrix orientation (Q1).  Arrays: c[s], m[s*s] row-major as used by the
 * stage loop, b[s], e[s].  Returns the number of stages s. */
int32_t or_rk_tableau(int32_t method, int32_t semantics, double *c, double *m,
                      double *b, double *e);
int32_t or_bdf_coefficients(int32_t method, double *higher, double *lower);
/* Adams predictor / corrector coefficients and the error coefficient
 * (adams.rs:568-661); returns the order O (5 or 3). */
int32_t or_adams_coefficients(int32_t method, double *pred, double *corr, double *err);
/* x^(1/n), n = 3 or 5, the DEVICE profile's exact-operation root */
double or_rootn(double x, int32_t n);
/* the DEVICE profile's step-factor root r^(1/8) (before the tangent bound) */
double or_root8(double r);

/* Discrete adjoint of the accepted RK45 (intended) step sequence and a frozen
 * step replay, used to check the GPU vjp.  See oracle/bacon_oracle.c. */
int32_t or_rk_replay(const or_problem *p, int64_t n_steps, const double *t_seq,
                     const double *dt_seq, double *y_out);

#ifdef __cplusplus
}
#endif
#endif
