/*
 * bacon_oracle.c — CPU restatement of bacon-sci 0.16.2's adaptive IVP step
 * loop.  TEST INFRASTRUCTURE ONLY: the checker and CPU-baseline timer for the
 * HIP product in bacon_amd/.  Nothing in the product loads this file.
 *
 * Followed line by line (paths relative to the bacon-sci crate root):
 *   src/ivp.rs:20-239       IVPStatus, IVPError order, IVPIterator::next
 *   src/ivp/rk.rs:168-247   builder setters and their validation
 *   src/ivp/rk.rs:249-343   RungeKutta::solve (constants, dt0, zeroed stages)
 *   src/ivp/rk.rs:361-423   RungeKuttaSolver::step (the RK hot loop)
 *   src/ivp/rk.rs:430-526   Fehlberg 4(5) coefficients
 *   src/ivp/rk.rs:563-621   Bogacki-Shampine 2(3) coefficients
 *   src/ivp/bdf.rs:257-331  BDF::solve
 *   src/ivp/bdf.rs:346-387  RK4 bootstrap
 *   src/ivp/bdf.rs:390-411  finite-difference Jacobian
 *   src/ivp/bdf.rs:414-475  Broyden ("secant") corrector
 *   src/ivp/bdf.rs:495-634  BDFSolver::step
 *   src/ivp/bdf.rs:641-673, 708-730  BDF6 / BDF2 coefficients
 * and the semantics of the third-party nalgebra 0.32 routines those lines
 * call (Matrix::from_vec column-major fill, norm() via dotc's 8-way unrolled
 * accumulation, gemv column-axpy order, LU / full-pivot LU / QR try_inverse).
 * nalgebra is not vendored with the reference; its behaviour is restated from
 * its published algorithms and the exact rounding of those calls is
 * "parity unpinned" (see the tests under tests/ that pin what can be pinned).
 *
 * Quirks of the reference (column-major tableau Q1, FD Jacobian sum Q3, lower
 * residual history Q4, f at t_n Q5, bootstrap time restore Q6) are data
 * switches selected by or_problem.semantics; no quirk is a separate path.
 *
 * Compile with -ffp-contract=off: Rust never fuses a*b+c, so the BACON
 * profile must round every product and every sum.
 */
#include "bacon_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD OR_MAX_DIM
#define DEFAULT_MAX_ATTEMPTS 10000000LL

/* ------------------------------------------------------------------ */
/* arithmetic profile helpers                                          */
/* ------------------------------------------------------------------ */

/* a*b + c: two roundings for BACON, one (fma) for DEVICE. */
static inline double madd(int prof, double a, double b, double c) {
    return prof == OR_PROFILE_DEVICE ? fma(a, b, c) : a * b + c;
}

/* nalgebra Matrix::norm() for a column vector: sqrt(dotc(x, x)).  dotc sums
 * sequentially below 8 rows and with eight interleaved accumulators above
 * (nalgebra 0.32 base/blas.rs dotx). */
static double nal_norm(int n, const double *x) {
    double res = 0.0;
    double acc[8];
    int i = 0;
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    while (n - i >= 8) {
        for (int k = 0; k < 8; ++k) acc[k] += x[i + k] * x[i + k];
        i += 8;
    }
    res += acc[0] + acc[4];
    res += acc[1] + acc[5];
    res += acc[2] + acc[6];
    res += acc[3] + acc[7];
    for (; i < n; ++i) res += x[i] * x[i];
    return sqrt(res);
}

/* nalgebra's norm() for a column vector of nc complex entries stored as
 * (re, im) pairs: dotc(x, x) sums conj(x_i) x_i, whose real part is
 * re_i re_i - (-im_i) im_i, one term per complex entry, with dotc's
 * accumulation order over the entries (sequential below 8, eight
 * interleaved accumulators above). */
static double nal_norm_c(int nc, const double *x) {
    double res = 0.0;
    double acc[8];
    int i = 0;
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    while (nc - i >= 8) {
        for (int k = 0; k < 8; ++k) {
            const double re = x[2 * (i + k)], im = x[2 * (i + k) + 1];
            acc[k] += re * re - (-im) * im;
        }
        i += 8;
    }
    res += acc[0] + acc[4];
    res += acc[1] + acc[5];
    res += acc[2] + acc[6];
    res += acc[3] + acc[7];
    for (; i < nc; ++i) {
        const double re = x[2 * i], im = x[2 * i + 1];
        res += re * re - (-im) * im;
    }
    return sqrt(res);
}

/* Device-order sum of squares: sequential fused accumulation. */
static double dev_sumsq(int n, const double *x) {
    double s = x[0] * x[0];
    for (int i = 1; i < n; ++i) s = fma(x[i], x[i], s);
    return s;
}

/* r^(1/8) for r in (1e-8, 65536), built from exact operations so that CPU and
 * GPU agree bit for bit: r = m 2^e with m in [1,2) (frexp), m^(1/8) from a
 * degree-5 polynomial in (m - 1.5) (Chebyshev-node interpolant, relative
 * error below 1.11e-6 over [1, 2)), 2^(e/8) = 2^floor(e/8) * 2^((e mod 8)/8) with the second
 * factor a product of 2^(1/8), 2^(1/4), 2^(1/2) selected by the bits of
 * e mod 8.  A NaN argument returns NaN.  The caller bounds the result by the
 * tangent at r = 1, which decides the controller near r = 1. */
static double dev_root8(double r) {
    static const double C[6] = {1.0519887150881322, 0.08766623673998887,
                                -0.025512534910039072, 0.010621966427993313,
                                -0.005693672682686368, 0.0029682975037312034};
    int ex = 0;
    const double mant = frexp(r, &ex); /* r = mant 2^ex, mant in [0.5, 1) */
    const double x = mant * 2.0 - 1.5;
    const int e = ex - 1;
    double p = C[5];
    for (int k = 4; k >= 0; --k) p = fma(p, x, C[k]);
    const int a = e >> 3, b = e & 7;
    const double s1 = (b & 1) ? 1.0905077326652577 : 1.0;
    const double s2 = (b & 2) ? 1.189207115002721 : 1.0;
    const double s4 = (b & 4) ? 1.4142135623730951 : 1.0;
    return ldexp(p * ((s1 * s2) * s4), a);
}

/* ------------------------------------------------------------------ */
/* right-hand sides                                                    */
/* ------------------------------------------------------------------ */

/* Complex state (bacon-sci's N: ComplexField, rk.rs:53, adams.rs:44): with
 * real times and step sizes, the solvers' complex products by real
 * coefficients act on the real and imaginary parts separately, so a system
 * of m complex components runs as 2m reals stored (re, im); only the error
 * norm groups each pair (nal_norm_c). */
int32_t or_rhs_is_complex(int32_t rhs) { return rhs == OR_RHS_TWO_LEVEL; }

double or_norm(int32_t n, const double *x, int32_t profile, int32_t complex_pairs) {
    if (n < 1) return 0.0;
    if (profile == OR_PROFILE_DEVICE) return sqrt(dev_sumsq(n, x));
    return complex_pairs ? nal_norm_c(n / 2, x) : nal_norm(n, x);
}

int32_t or_rhs_eval(int32_t rhs, int32_t prof, int32_t dim, double t,
                    const double *y, const double *pr, const double *shared,
                    double *f) {
    switch (rhs) {
    case OR_RHS_EXP:
        for (int i = 0; i < dim; ++i) f[i] = y[i];
        return OR_OK;
    case OR_RHS_NEG:
        for (int i = 0; i < dim; ++i) f[i] = -y[i];
        return OR_OK;
    case OR_RHS_QUADRATIC:
        for (int i = 0; i < dim; ++i) f[i] = -2.0 * t;
        return OR_OK;
    case OR_RHS_COS:
        for (int i = 0; i < dim; ++i) f[i] = cos(t);
        return OR_OK;
    case OR_RHS_LORENZ: {
        if (dim != 3) return OR_INVALID_ARGUMENT;
        const double s = pr[0], r = pr[1], b = pr[2];
        f[0] = s * (y[1] - y[0]);
        if (prof == OR_PROFILE_DEVICE) {
            f[1] = fma(y[0], r - y[2], -y[1]);
            f[2] = fma(y[0], y[1], -(b * y[2]));
        } else {
            f[1] = y[0] * (r - y[2]) - y[1];
            f[2] = y[0] * y[1] - b * y[2];
        }
        return OR_OK;
    }
    case OR_RHS_LINEAR: {
        const double *A = shared, *bv = shared + (size_t)dim * dim;
        for (int i = 0; i < dim; ++i) {
            double acc = bv[i];
            for (int j = 0; j < dim; ++j)
                acc = madd(prof, A[(size_t)i * dim + j], y[j], acc);
            f[i] = acc;
        }
        return OR_OK;
    }
    case OR_RHS_BRUSSELATOR: {
        /* y = [u_0..u_{n-1}, v_0..v_{n-1}], periodic, dx = 1/n */
        if (dim % 2) return OR_INVALID_ARGUMENT;
        const int n = dim / 2;
        const double A = pr[0], B = pr[1], alpha = pr[2];
        const double dx = 1.0 / n;
        const double c = alpha / (dx * dx);
        const double *u = y, *v = y + n;
        for (int i = 0; i < n; ++i) {
            const int im = (i + n - 1) % n, ip = (i + 1) % n;
            const double uuv = u[i] * u[i] * v[i];
            const double lu = u[im] - 2.0 * u[i] + u[ip];
            const double lv = v[im] - 2.0 * v[i] + v[ip];
            f[i] = A + uuv - (B + 1.0) * u[i] + c * lu;
            f[n + i] = B * u[i] - uuv + c * lv;
        }
        return OR_OK;
    }
    case OR_RHS_LOTKA_VOLTERRA: {
        if (dim != 2) return OR_INVALID_ARGUMENT;
        const double a = pr[0], b = pr[1], g = pr[2], d = pr[3];
        const double xy = y[0] * y[1];
        if (prof == OR_PROFILE_DEVICE) {
            f[0] = fma(a, y[0], -(b * xy));
            f[1] = fma(d, xy, -(g * y[1]));
        } else {
            f[0] = a * y[0] - b * xy;
            f[1] = d * xy - g * y[1];
        }
        return OR_OK;
    }
    case OR_RHS_OSCILLATOR:
        if (dim != 2) return OR_INVALID_ARGUMENT;
        f[0] = y[1];
        f[1] = -y[0];
        return OR_OK;
    case OR_RHS_ROBERTSON: {
        /* Robertson's kinetics; separate products and sums in both profiles */
        if (dim != 3) return OR_INVALID_ARGUMENT;
        const double a = pr[0] * y[0];
        const double b = pr[2] * (y[1] * y[2]);
        const double c = pr[1] * (y[1] * y[1]);
        f[0] = b - a;
        f[1] = (a - b) - c;
        f[2] = c;
        return OR_OK;
    }
    case OR_RHS_TWO_LEVEL: {
        /* y = (a0, b0, a1, b1), components c_k = a_k + i b_k;
         * -i H c = H b - i H a; then -gamma c */
        if (dim != 4) return OR_INVALID_ARGUMENT;
        const double hd = pr[0], ho = pr[1], g = pr[2];
        double ha0, ha1, hb0, hb1;
        if (prof == OR_PROFILE_DEVICE) {
            ha0 = fma(hd, y[0], ho * y[2]);
            ha1 = fma(ho, y[0], -(hd * y[2]));
            hb0 = fma(hd, y[1], ho * y[3]);
            hb1 = fma(ho, y[1], -(hd * y[3]));
            f[0] = fma(-g, y[0], hb0);
            f[1] = fma(-g, y[1], -ha0);
            f[2] = fma(-g, y[2], hb1);
            f[3] = fma(-g, y[3], -ha1);
        } else {
            ha0 = hd * y[0] + ho * y[2];
            ha1 = ho * y[0] - hd * y[2];
            hb0 = hd * y[1] + ho * y[3];
            hb1 = ho * y[1] - hd * y[3];
            f[0] = hb0 - g * y[0];
            f[1] = -ha0 - g * y[1];
            f[2] = hb1 - g * y[2];
            f[3] = -ha1 - g * y[3];
        }
        return OR_OK;
    }
    case OR_RHS_FAIL_AFTER:
        if (t >= pr[0]) return OR_USER_ERROR;
        for (int i = 0; i < dim; ++i) f[i] = y[i];
        return OR_OK;
    case OR_RHS_LV_SENS: {
        /* z = [x, y, S(2x6) column-major]; columns: d/d(alpha, beta, gamma,
         * delta, x0, y0).  S' = J S + df/dp. */
        if (dim != 14) return OR_INVALID_ARGUMENT;
        const double a = pr[0], b = pr[1], g = pr[2], d = pr[3];
        const double x = y[0], w = y[1];
        const double xy = x * w;
        f[0] = a * x - b * xy;
        f[1] = d * xy - g * w;
        const double j00 = a - b * w, j01 = -b * x, j10 = d * w, j11 = d * x - g;
        const double fp[6][2] = {{x, 0.0}, {-xy, 0.0}, {0.0, -w},
                                 {0.0, xy}, {0.0, 0.0}, {0.0, 0.0}};
        for (int k = 0; k < 6; ++k) {
            const double s0 = y[2 + 2 * k], s1 = y[3 + 2 * k];
            f[2 + 2 * k] = j00 * s0 + j01 * s1 + fp[k][0];
            f[3 + 2 * k] = j10 * s0 + j11 * s1 + fp[k][1];
        }
        return OR_OK;
    }
    default:
        return OR_INVALID_ARGUMENT;
    }
}

/* ------------------------------------------------------------------ */
/* Runge-Kutta tableaux (rk.rs:430-526, 563-621)                       */
/* ------------------------------------------------------------------ */

typedef struct {
    int s;
    double c[6], m[6][6], b[6], e[6];
} rk_tab;

/* Build the tableau the stage loop actually uses.  The reference writes the
 * Butcher rows into Matrix::from_vec, which fills column-major, and then
 * iterates row_iter (rk.rs:370): the stage loop reads the TRANSPOSE of the
 * literal rows.  AS_BUILT keeps that; INTENDED reads the rows as rows. */
static int rk_build(int method, int semantics, rk_tab *t) {
    double lit[6][6];
    memset(t, 0, sizeof *t);
    memset(lit, 0, sizeof lit);
    if (method == OR_RK45) {
        t->s = 6;
        t->c[0] = 0.0;
        t->c[1] = 1.0 / 4.0; /* from_u8(4).recip() */
        t->c[2] = 3.0 / 8.0;
        t->c[3] = 12.0 / 13.0;
        t->c[4] = 1.0;
        t->c[5] = 1.0 / 2.0;
        lit[1][0] = 1.0 / 4.0;
        lit[2][0] = 3.0 / 32.0;
        lit[2][1] = 9.0 / 32.0;
        lit[3][0] = 1932.0 / 2197.0;
        lit[3][1] = -7200.0 / 2197.0;
        lit[3][2] = 7296.0 / 2197.0;
        lit[4][0] = 439.0 / 216.0;
        lit[4][1] = -8.0;
        lit[4][2] = 3680.0 / 513.0;
        lit[4][3] = -845.0 / 4104.0;
        lit[5][0] = -8.0 / 27.0;
        lit[5][1] = 2.0;
        lit[5][2] = -3544.0 / 2565.0;
        lit[5][3] = 1859.0 / 4104.0;
        lit[5][4] = -11.0 / 40.0;
        t->b[0] = 25.0 / 216.0;
        t->b[1] = 0.0;
        t->b[2] = 1408.0 / 2565.0;
        t->b[3] = 2197.0 / 4104.0;
        t->b[4] = -(1.0 / 5.0);
        t->b[5] = 0.0;
        t->e[0] = 1.0 / 360.0;
        t->e[1] = 0.0;
        t->e[2] = -128.0 / 4275.0;
        t->e[3] = -2197.0 / 75240.0;
        t->e[4] = 1.0 / 50.0;
        t->e[5] = 2.0 / 55.0;
    } else if (method == OR_RK23) {
        t->s = 4;
        t->c[0] = 0.0;
        t->c[1] = 1.0 / 2.0;
        t->c[2] = 3.0 / 4.0;
        t->c[3] = 1.0;
        lit[1][0] = 1.0 / 2.0;
        lit[2][1] = 3.0 / 4.0;
        lit[3][0] = 2.0 / 9.0;
        lit[3][1] = 1.0 / 3.0;
        lit[3][2] = 4.0 / 9.0;
        t->b[0] = 2.0 / 9.0;
        t->b[1] = 1.0 / 3.0;
        t->b[2] = 4.0 / 9.0;
        t->b[3] = 0.0;
        t->e[0] = -5.0 / 72.0;
        t->e[1] = 1.0 / 12.0;
        t->e[2] = 1.0 / 9.0;
        t->e[3] = -(1.0 / 8.0);
    } else {
        return OR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < t->s; ++i)
        for (int j = 0; j < t->s; ++j)
            t->m[i][j] = semantics == OR_AS_BUILT ? lit[j][i] : lit[i][j];
    return OR_OK;
}

int32_t or_rk_tableau(int32_t method, int32_t semantics, double *c, double *m,
                      double *b, double *e) {
    rk_tab t;
    if (rk_build(method, semantics, &t) != OR_OK) return OR_INVALID_ARGUMENT;
    for (int i = 0; i < t.s; ++i) {
        c[i] = t.c[i];
        b[i] = t.b[i];
        e[i] = t.e[i];
        for (int j = 0; j < t.s; ++j) m[i * t.s + j] = t.m[i][j];
    }
    return t.s;
}

/* ------------------------------------------------------------------ */
/* RK solver (rk.rs:249-423)                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    const or_problem *p;
    rk_tab tab;
    int prof, D;
    double dt_max, dt_min, time, end, tol, dt;
    double state[MAXD];
    double K[6][MAXD]; /* half_steps: column i = stage i, persists (Q1) */
    double one_tenth, one_fourth, point_eighty_four, four;
    int next_stop;
    double *stop_y;
    int64_t n_rhs;
} rk_solver;

static int rk_init(rk_solver *s, const or_problem *p, double *stop_y) {
    memset(s, 0, sizeof *s);
    s->p = p;
    if (rk_build(p->method, p->semantics, &s->tab) != OR_OK)
        return OR_INVALID_ARGUMENT;
    s->prof = p->profile;
    s->D = p->dim;
    s->dt_max = p->dt_max;
    s->dt_min = p->dt_min;
    s->time = p->t0;
    s->end = p->t_end;
    s->tol = p->tol;
    const double half = 1.0 / 2.0;
    s->dt = (p->dt_max + p->dt_min) * half; /* rk.rs:315 */
    memcpy(s->state, p->y0, sizeof(double) * (size_t)p->dim);
    s->one_tenth = 1.0 / 10.0;
    s->four = 4.0;
    s->one_fourth = 1.0 / s->four;
    /* rk.rs:266-268: both constants are from_u8(100), so the safety factor
     * is 100/100 = 1.0 (Q2; kept in both semantics). */
    s->point_eighty_four = 100.0 / 100.0;
    s->stop_y = stop_y;
    return OR_OK;
}

/* One call of RungeKuttaSolver::step.  Returns OR_STEP_YIELD (Ok),
 * OR_STEP_REDO, OR_STEP_DONE, or an IVPError code.  *used_dt receives the
 * step size of this attempt. */
static int rk_step(rk_solver *s, double *used_dt) {
    const int D = s->D, S = s->tab.s, prof = s->prof;
    const or_problem *p = s->p;
    double scratch[MAXD], f[MAXD];

    if (s->time >= s->end) return OR_STEP_DONE; /* rk.rs:362-364 */

    /* rk.rs:366-368, generalised to observation stops strictly inside the
     * interval (an extension; with n_stops == 0 it is the reference clamp). */
    double target = s->end;
    int at_stop = 0;
    if (s->next_stop < p->n_stops) {
        target = p->stops[s->next_stop];
        at_stop = 1;
    }
    int clamped = 0;
    const double dt_proposed = s->dt;
    if (s->time + s->dt >= target) {
        s->dt = target - s->time;
        clamped = 1;
    }
    *used_dt = s->dt;

    /* rk.rs:370-384 */
    for (int i = 0; i < S; ++i) {
        memcpy(scratch, s->state, sizeof(double) * (size_t)D);
        for (int j = 0; j < S; ++j) {
            const double m = s->tab.m[i][j];
            if (prof == OR_PROFILE_DEVICE) {
                if (m == 0.0) continue;
                for (int d = 0; d < D; ++d) scratch[d] = fma(m, s->K[j][d], scratch[d]);
            } else {
                for (int d = 0; d < D; ++d) scratch[d] += s->K[j][d] * m;
            }
        }
        const double step_time = prof == OR_PROFILE_DEVICE
                                     ? fma(s->tab.c[i], s->dt, s->time)
                                     : s->time + s->tab.c[i] * s->dt;
        int rc = or_rhs_eval(p->rhs, prof, D, step_time, scratch, p->params,
                             p->shared, f);
        s->n_rhs++;
        if (rc != OR_OK) return rc == OR_INVALID_ARGUMENT ? rc : OR_USER_ERROR;
        for (int d = 0; d < D; ++d) s->K[i][d] = f[d] * s->dt;
    }

    /* rk.rs:386-390 */
    if (prof == OR_PROFILE_DEVICE) {
        int first = 1;
        for (int j = 0; j < S; ++j) {
            const double e = s->tab.e[j];
            if (e == 0.0) continue;
            for (int d = 0; d < D; ++d)
                scratch[d] = first ? e * s->K[j][d] : fma(e, s->K[j][d], scratch[d]);
            first = 0;
        }
    } else {
        for (int d = 0; d < D; ++d) scratch[d] = s->K[0][d] * s->tab.e[0];
        for (int j = 1; j < S; ++j)
            for (int d = 0; d < D; ++d) scratch[d] += s->K[j][d] * s->tab.e[j];
    }
    /* rk.rs:390-412.  BACON: err = |E|/dt, accept err <= tol,
     * delta = 1.0 * (tol/err)^(1/4), clamps 0.1 / 4.  DEVICE: the same
     * decisions from ss = |E|^2 and q = tol*dt: accept ss <= q*q, and
     * delta = (q*q/ss)^(1/8) with the clamps applied to r = q*q/ss
     * (delta <= 0.1 <=> r <= 1e-8, delta >= 4 <=> r >= 65536); the root is
     * dev_root8 above.  Both profiles keep the reference's branch structure. */
    int accepted;
    double factor;
    if (prof == OR_PROFILE_DEVICE) {
        const double ss = dev_sumsq(D, scratch);
        const double q = s->tol * s->dt;
        const double q2 = q * q;
        accepted = ss <= q2;
        const double r = q2 / ss;
        const double f0 = r <= 1e-8 ? s->one_tenth : (r >= 65536.0 ? s->four : dev_root8(r));
        /* Near r = 1 the controller's decisions hinge on the sign of
         * delta - 1 = (r - 1)/8 + O((r-1)^2), far below the polynomial's
         * 1.1e-6 relative error.  r^(1/8) is concave, so its tangent at 1,
         * 1 + (r - 1)/8, bounds it from above and is exact to second order
         * there: take the smaller of the two.  Then keep the exact power's
         * side of 1 (delta < 1 on a rejection, >= 1 on an acceptance) against
         * the final rounding.  fmin/fmax (IEEE minNum/maxNum) are single
         * instructions on the device; they differ from compare-and-select only
         * for NaN: a NaN error estimate (rejected, like the reference) gets the
         * factor 1 - 2^-53 instead of NaN, so dt stays finite and the instance
         * ends at the attempt cap either way (the reference retries forever). */
        const double tang = fma(r - 1.0, 0.125, 1.0);
        const double f = fmin(f0, tang);
        factor = accepted ? fmax(f, 1.0) : fmin(f, 0x1.fffffffffffffp-1);
    } else {
        const double nrm = or_rhs_is_complex(p->rhs) ? nal_norm_c(D / 2, scratch)
                                                      : nal_norm(D, scratch);
        const double error = nrm / s->dt;
        accepted = error <= s->tol;
        const double delta =
            s->point_eighty_four * pow(s->tol / error, s->one_fourth);
        factor = delta <= s->one_tenth ? s->one_tenth : (delta >= s->four ? s->four : delta);
    }

    /* rk.rs:392-398: the 4th-order weights propagate the state */
    if (accepted) {
        s->time += s->dt;
        for (int j = 0; j < S; ++j) {
            const double b = s->tab.b[j];
            if (prof == OR_PROFILE_DEVICE) {
                if (b == 0.0) continue;
                for (int d = 0; d < D; ++d) s->state[d] = fma(b, s->K[j][d], s->state[d]);
            } else {
                for (int d = 0; d < D; ++d) s->state[d] += s->K[j][d] * b;
            }
        }
        if (clamped && at_stop) {
            /* land exactly on the observation time and record it */
            s->time = target;
            if (s->stop_y)
                memcpy(s->stop_y + (size_t)s->next_stop * D, s->state,
                       sizeof(double) * (size_t)D);
            s->next_stop++;
        }
    }

    /* rk.rs:400-412 */
    s->dt *= factor;
    if (s->dt > s->dt_max) s->dt = s->dt_max;
    /* Extension for stops: an accepted step cut short to land on a stop hands
     * the next attempt at least the step size proposed before the cut, so a
     * stop just past an accepted time cannot push dt below dt_min (without
     * stops the clamp only ever lands on the end, where the solve is done). */
    if (accepted && clamped && at_stop && s->dt < dt_proposed) s->dt = dt_proposed;

    /* rk.rs:414-416 (Q8: an accepted step can still fail here) */
    if (s->dt < s->dt_min && s->time < s->end) return OR_MINIMUM_TIME_DELTA_EXCEEDED;

    return accepted ? OR_STEP_YIELD : OR_STEP_REDO;
}

static void path_push(or_path *path, int D, double t, const double *y, double dt) {
    if (!path) return;
    const int64_t k = path->len++;
    if (k >= path->cap) return;
    if (path->t) path->t[k] = t;
    if (path->dt) path->dt[k] = dt;
    if (path->y) memcpy(path->y + (size_t)k * D, y, sizeof(double) * (size_t)D);
}

static int rk_solve(const or_problem *p, double *y_out, or_result *r,
                    or_path *path, double *stop_y) {
    rk_solver s;
    int rc = rk_init(&s, p, stop_y);
    if (rc != OR_OK) return rc;
    const int64_t cap = p->max_attempts > 0 ? p->max_attempts : DEFAULT_MAX_ATTEMPTS;
    int64_t attempts = 0;
    int status = OR_OK;
    /* IVPIterator::next (ivp.rs:220-238): Ok yields, Redo retries, Done ends,
     * Failure ends the iteration. */
    for (;;) {
        if (attempts >= cap) {
            /* the reference would spin forever on a NaN error (Q9); the
             * attempt cap turns that into a reported failure */
            status = OR_MAXIMUM_ITERATIONS_EXCEEDED;
            break;
        }
        double used_dt = 0.0;
        const int st = rk_step(&s, &used_dt);
        if (st == OR_STEP_DONE) break;
        attempts++;
        if (st == OR_STEP_YIELD) {
            r->n_accepted++;
            r->n_yield++;
            path_push(path, s.D, s.time, s.state, used_dt);
        } else if (st == OR_STEP_REDO) {
            r->n_rejected++;
        } else {
            status = st;
            break;
        }
    }
    r->t = s.time;
    r->dt = s.dt;
    r->status = status;
    r->n_rhs = s.n_rhs;
    memcpy(y_out, s.state, sizeof(double) * (size_t)s.D);
    return status;
}

/* Frozen-step replay: apply the accepted RK step map with the given (t, dt)
 * sequence, no controller.  y_out receives n_steps*dim states. */
int32_t or_rk_replay(const or_problem *p, int64_t n_steps, const double *t_seq,
                     const double *dt_seq, double *y_out) {
    rk_tab tab;
    if (rk_build(p->method, p->semantics, &tab) != OR_OK) return OR_INVALID_ARGUMENT;
    const int D = p->dim, S = tab.s, prof = p->profile;
    double y[MAXD], K[6][MAXD], sc[MAXD], f[MAXD];
    memcpy(y, p->y0, sizeof(double) * (size_t)D);
    memset(K, 0, sizeof K);
    for (int64_t n = 0; n < n_steps; ++n) {
        const double t = t_seq[n], dt = dt_seq[n];
        for (int i = 0; i < S; ++i) {
            memcpy(sc, y, sizeof(double) * (size_t)D);
            for (int j = 0; j < S; ++j) {
                const double m = tab.m[i][j];
                if (m == 0.0) continue;
                for (int d = 0; d < D; ++d) sc[d] = madd(prof, m, K[j][d], sc[d]);
            }
            const double ti = madd(prof, tab.c[i], dt, t);
            int rc = or_rhs_eval(p->rhs, prof, D, ti, sc, p->params, p->shared, f);
            if (rc != OR_OK) return rc;
            for (int d = 0; d < D; ++d) K[i][d] = f[d] * dt;
        }
        for (int j = 0; j < S; ++j) {
            if (tab.b[j] == 0.0) continue;
            for (int d = 0; d < D; ++d) y[d] = madd(prof, tab.b[j], K[j][d], y[d]);
        }
        memcpy(y_out + (size_t)n * D, y, sizeof(double) * (size_t)D);
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* nalgebra dense inverse cascade (bdf.rs:429-444)                     */
/* column-major storage: a[i + j*n]                                    */
/* ------------------------------------------------------------------ */

#define AT(a, n, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(n)]

/* lu::gauss_step / gauss_step_swap: scale the sub-diagonal of column i by
 * 1/diag, then rank-1 update of the trailing block column by column. */
static void gauss_step(int n, double *m, double diag, int i, int piv) {
    const double inv_diag = 1.0 / diag;
    if (piv != i) {
        double tmp = AT(m, n, i, i);
        AT(m, n, i, i) = AT(m, n, piv, i);
        AT(m, n, piv, i) = tmp;
    }
    for (int r = i + 1; r < n; ++r) AT(m, n, r, i) *= inv_diag;
    for (int k = i + 1; k < n; ++k) {
        if (piv != i) {
            double tmp = AT(m, n, i, k);
            AT(m, n, i, k) = AT(m, n, piv, k);
            AT(m, n, piv, k) = tmp;
        }
        const double a = -AT(m, n, i, k);
        for (int r = i + 1; r < n; ++r)
            AT(m, n, r, k) = a * AT(m, n, r, i) + AT(m, n, r, k);
    }
}

/* triangular solves of nalgebra solve.rs, column by column */
static void lower_unit_solve(int n, const double *lu, double *b) {
    for (int i = 0; i < n - 1; ++i) {
        const double coeff = b[i] / 1.0;
        for (int r = i + 1; r < n; ++r) b[r] = (-coeff) * AT(lu, n, r, i) + b[r];
    }
}

static int upper_solve(int n, const double *lu, double *b) {
    for (int i = n - 1; i >= 0; --i) {
        const double diag = AT(lu, n, i, i);
        if (diag == 0.0) return 0;
        const double coeff = b[i] / diag;
        b[i] = coeff;
        for (int r = 0; r < i; ++r) b[r] = (-coeff) * AT(lu, n, r, i) + b[r];
    }
    return 1;
}

int32_t or_lu_inverse(int32_t n, const double *a, double *inv) {
    double *m = (double *)malloc(sizeof(double) * (size_t)n * n);
    int *perm_i = (int *)malloc(sizeof(int) * (size_t)n);
    int *perm_p = (int *)malloc(sizeof(int) * (size_t)n);
    int np = 0, ok = 1;
    memcpy(m, a, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        /* icamax: first index of the largest |x| */
        int piv = i;
        double best = fabs(AT(m, n, i, i));
        for (int r = i + 1; r < n; ++r) {
            const double v = fabs(AT(m, n, r, i));
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        const double diag = AT(m, n, piv, i);
        if (diag == 0.0) continue;
        if (piv != i) {
            perm_i[np] = i;
            perm_p[np] = piv;
            np++;
            for (int k = 0; k < i; ++k) {
                double tmp = AT(m, n, i, k);
                AT(m, n, i, k) = AT(m, n, piv, k);
                AT(m, n, piv, k) = tmp;
            }
        }
        gauss_step(n, m, diag, i, piv);
    }
    double *col = (double *)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < n && ok; ++j) {
        for (int r = 0; r < n; ++r) col[r] = (r == j) ? 1.0 : 0.0;
        for (int k = 0; k < np; ++k) {
            double tmp = col[perm_i[k]];
            col[perm_i[k]] = col[perm_p[k]];
            col[perm_p[k]] = tmp;
        }
        lower_unit_solve(n, m, col);
        if (!upper_solve(n, m, col)) ok = 0;
        for (int r = 0; r < n; ++r) AT(inv, n, r, j) = col[r];
    }
    free(col);
    free(m);
    free(perm_i);
    free(perm_p);
    return ok ? OR_OK : OR_SINGULAR_MATRIX;
}

/* FullPivLU: pivot on the largest |x| of the trailing block (column-major
 * scan, first maximum), stop at an exact zero pivot. */
static int full_piv_inverse(int n, const double *a, double *inv) {
    double *m = (double *)malloc(sizeof(double) * (size_t)n * n);
    int *pi = (int *)malloc(sizeof(int) * (size_t)n), *pp = (int *)malloc(sizeof(int) * (size_t)n);
    int *qi = (int *)malloc(sizeof(int) * (size_t)n), *qp = (int *)malloc(sizeof(int) * (size_t)n);
    int np = 0, nq = 0;
    memcpy(m, a, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        int br = i, bc = i;
        double best = fabs(AT(m, n, i, i));
        for (int c = i; c < n; ++c)
            for (int r = i; r < n; ++r) {
                const double v = fabs(AT(m, n, r, c));
                if (v > best) {
                    best = v;
                    br = r;
                    bc = c;
                }
            }
        const double diag = AT(m, n, br, bc);
        if (diag == 0.0) break;
        if (bc != i)
            for (int r = 0; r < n; ++r) {
                double tmp = AT(m, n, r, i);
                AT(m, n, r, i) = AT(m, n, r, bc);
                AT(m, n, r, bc) = tmp;
            }
        qi[nq] = i;
        qp[nq] = bc;
        nq++;
        if (br != i) {
            pi[np] = i;
            pp[np] = br;
            np++;
            for (int k = 0; k < i; ++k) {
                double tmp = AT(m, n, i, k);
                AT(m, n, i, k) = AT(m, n, br, k);
                AT(m, n, br, k) = tmp;
            }
        }
        gauss_step(n, m, diag, i, br);
    }
    int ok = AT(m, n, n - 1, n - 1) != 0.0;
    double *col = (double *)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < n && ok; ++j) {
        for (int r = 0; r < n; ++r) col[r] = (r == j) ? 1.0 : 0.0;
        for (int k = 0; k < np; ++k) {
            double tmp = col[pi[k]];
            col[pi[k]] = col[pp[k]];
            col[pp[k]] = tmp;
        }
        lower_unit_solve(n, m, col);
        upper_solve(n, m, col);
        for (int k = nq - 1; k >= 0; --k) { /* inverse column permutation */
            double tmp = col[qi[k]];
            col[qi[k]] = col[qp[k]];
            col[qp[k]] = tmp;
        }
        for (int r = 0; r < n; ++r) AT(inv, n, r, j) = col[r];
    }
    free(col);
    free(m);
    free(pi);
    free(pp);
    free(qi);
    free(qp);
    return ok;
}

/* Householder QR inverse; fails on an exact zero diagonal of R. */
static int qr_inverse(int n, const double *a, double *inv) {
    double *m = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *v = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *beta = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(m, a, sizeof(double) * (size_t)n * n);
    for (int k = 0; k < n; ++k) {
        double nrm = 0.0;
        for (int r = k; r < n; ++r) nrm += AT(m, n, r, k) * AT(m, n, r, k);
        nrm = sqrt(nrm);
        for (int r = 0; r < n; ++r) AT(v, n, r, k) = 0.0;
        beta[k] = 0.0;
        if (nrm == 0.0) continue;
        const double alpha = AT(m, n, k, k) >= 0.0 ? -nrm : nrm;
        for (int r = k; r < n; ++r) AT(v, n, r, k) = AT(m, n, r, k);
        AT(v, n, k, k) -= alpha;
        double vv = 0.0;
        for (int r = k; r < n; ++r) vv += AT(v, n, r, k) * AT(v, n, r, k);
        if (vv == 0.0) continue;
        beta[k] = 2.0 / vv;
        for (int c = k; c < n; ++c) {
            double dot = 0.0;
            for (int r = k; r < n; ++r) dot += AT(v, n, r, k) * AT(m, n, r, c);
            dot *= beta[k];
            for (int r = k; r < n; ++r) AT(m, n, r, c) -= dot * AT(v, n, r, k);
        }
    }
    int ok = 1;
    for (int k = 0; k < n; ++k)
        if (AT(m, n, k, k) == 0.0) ok = 0;
    double *col = (double *)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < n && ok; ++j) {
        for (int r = 0; r < n; ++r) col[r] = (r == j) ? 1.0 : 0.0;
        for (int k = 0; k < n; ++k) { /* col = Q^T e_j */
            if (beta[k] == 0.0) continue;
            double dot = 0.0;
            for (int r = k; r < n; ++r) dot += AT(v, n, r, k) * col[r];
            dot *= beta[k];
            for (int r = k; r < n; ++r) col[r] -= dot * AT(v, n, r, k);
        }
        upper_solve(n, m, col);
        for (int r = 0; r < n; ++r) AT(inv, n, r, j) = col[r];
    }
    free(col);
    free(m);
    free(v);
    free(beta);
    return ok;
}

/* One member of the cascade on its own (for the tests that pin each against
 * LAPACK): branch 0 = partial-pivot LU, 1 = full-pivot LU, 2 = QR. */
int32_t or_inverse_branch(int32_t n, const double *a, double *inv, int32_t branch) {
    switch (branch) {
    case 0:
        return or_lu_inverse(n, a, inv);
    case 1:
        return full_piv_inverse(n, a, inv) ? OR_OK : OR_SINGULAR_MATRIX;
    case 2:
        return qr_inverse(n, a, inv) ? OR_OK : OR_SINGULAR_MATRIX;
    default:
        return OR_INVALID_ARGUMENT;
    }
}

int32_t or_inverse(int32_t n, const double *a, double *inv, int32_t *which) {
    if (or_lu_inverse(n, a, inv) == OR_OK) {
        if (which) *which = 0;
        return OR_OK;
    }
    if (full_piv_inverse(n, a, inv)) {
        if (which) *which = 1;
        return OR_OK;
    }
    if (qr_inverse(n, a, inv)) {
        if (which) *which = 2;
        return OR_OK;
    }
    if (which) *which = -1;
    return OR_SINGULAR_MATRIX;
}

/* ------------------------------------------------------------------ */
/* BDF solver (bdf.rs:257-634)                                         */
/* ------------------------------------------------------------------ */

#define HCAP 16

typedef struct {
    const or_problem *p;
    int O, D;
    double dt_max, dt_min, time, end, tol, dt;
    double state[MAXD], save_state[MAXD];
    double higher[7], lower[7];
    /* prev_values: VecDeque<(t, y)> as a ring buffer */
    double ht[HCAP];
    double hy[HCAP][MAXD];
    int hhead, hlen;
    double one_tenth, one_sixth, half, two, order;
    int yield_memory;
    /* semantics switches */
    double fd_sign;      /* +1 as built (sum, Q3), -1 intended         */
    int lower_own_hist;  /* 0 as built (Q4), 1 intended               */
    double f_time_theta; /* 0 as built (f at t_n, Q5), 1 intended      */
    int restore_mul;     /* 0 as built (t -= dt - O, Q6), 1: t -= O*dt */
    int64_t n_rhs, n_newton, n_lu, n_reject;
    double jinv[MAXD * MAXD], jac[MAXD * MAXD];
} bdf_solver;

int32_t or_bdf_coefficients(int32_t method, double *hi, double *lo) {
    if (method == OR_BDF6) {
        const double h = 147.0, l = 137.0;
        hi[0] = 60.0 / h; hi[1] = -360.0 / h; hi[2] = 450.0 / h; hi[3] = -400.0 / h;
        hi[4] = 225.0 / h; hi[5] = -72.0 / h; hi[6] = 10.0 / h;
        lo[0] = 60.0 / l; lo[1] = -300.0 / l; lo[2] = 300.0 / l; lo[3] = -200.0 / l;
        lo[4] = 75.0 / l; lo[5] = -12.0 / l; lo[6] = 0.0;
        return 7;
    }
    if (method == OR_BDF2) {
        const double three = 3.0;
        hi[0] = 2.0 / three; hi[1] = -4.0 / three; hi[2] = 1.0 / three;
        lo[0] = 1.0; lo[1] = -1.0; lo[2] = 0.0;
        return 3;
    }
    return OR_INVALID_ARGUMENT;
}

static double *hist_at(bdf_solver *s, int k) { return s->hy[(s->hhead + k) % HCAP]; }
static double hist_t(bdf_solver *s, int k) { return s->ht[(s->hhead + k) % HCAP]; }
static void hist_push(bdf_solver *s, double t, const double *y) {
    const int slot = (s->hhead + s->hlen) % HCAP;
    s->ht[slot] = t;
    memcpy(s->hy[slot], y, sizeof(double) * (size_t)s->D);
    s->hlen++;
}
static void hist_pop_front(bdf_solver *s) {
    if (s->hlen == 0) return;
    s->hhead = (s->hhead + 1) % HCAP;
    s->hlen--;
}

static int bdf_rhs(bdf_solver *s, double t, const double *y, double *f) {
    s->n_rhs++;
    int rc = or_rhs_eval(s->p->rhs, OR_PROFILE_BACON, s->D, t, y, s->p->params,
                         s->p->shared, f);
    if (rc == OR_INVALID_ARGUMENT) return rc;
    return rc == OR_OK ? OR_OK : OR_USER_ERROR;
}

/* bdf.rs:346-387 classical RK4, pushing states 2..n and the final one */
static int bdf_runge_kutta(bdf_solver *s, int iterations) {
    const int D = s->D;
    double k1[MAXD], k2[MAXD], k3[MAXD], k4[MAXD], mid[MAXD];
    for (int it = 0; it < iterations; ++it) {
        int rc = bdf_rhs(s, s->time, s->state, k1);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k1[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = s->state[d] + k1[d] * s->half;
        rc = bdf_rhs(s, s->time + s->half * s->dt, mid, k2);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k2[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = s->state[d] + k2[d] * s->half;
        rc = bdf_rhs(s, s->time + s->half * s->dt, mid, k3);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k3[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = s->state[d] + k3[d];
        rc = bdf_rhs(s, s->time + s->dt, mid, k4);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k4[d] *= s->dt;
        if (it != 0) hist_push(s, s->time, s->state);
        for (int d = 0; d < D; ++d)
            s->state[d] += (k1[d] + k2[d] * s->two + k3[d] * s->two + k4[d]) * s->one_sixth;
        s->time += s->dt;
    }
    hist_push(s, s->time, s->state);
    return OR_OK;
}

/* residual closures bdf.rs:548-575:
 * G(y) = (-f(t, y)) * dt * beta0 + sum_k coeff_k * prev[O-k] + y */
static int bdf_residual(bdf_solver *s, int lower, double t, const double *y, double *g) {
    const int D = s->D, O = s->O;
    double f[MAXD];
    int rc = bdf_rhs(s, t, y, f);
    if (rc) return rc;
    const double beta0 = lower ? s->lower[0] : s->higher[0];
    const double *hc = (lower && s->lower_own_hist) ? s->lower : s->higher;
    for (int d = 0; d < D; ++d) g[d] = (-f[d]) * s->dt * beta0;
    for (int ind = 1; ind < O; ++ind) {
        const double *py = hist_at(s, O - ind);
        for (int d = 0; d < D; ++d) g[d] += py[d] * hc[ind];
    }
    for (int d = 0; d < D; ++d) g[d] = g[d] + y[d];
    return OR_OK;
}

/* gemv: out = (-A) x, column-axpy order (nalgebra gemv) */
static void neg_gemv(int n, const double *A, const double *x, double *out) {
    for (int i = 0; i < n; ++i) out[i] = x[0] * (-AT(A, n, i, 0));
    for (int j = 1; j < n; ++j)
        for (int i = 0; i < n; ++i) out[i] = x[j] * (-AT(A, n, i, j)) + out[i];
}

/* bdf.rs:414-475 */
static int bdf_secant(bdf_solver *s, int lower, double *guess) {
    const int D = s->D;
    const double t = s->time + s->f_time_theta * s->dt;
    double F[MAXD], Fl[MAXD], shift[MAXD], adj[MAXD], u[MAXD], above[MAXD], below[MAXD];
    int n = 2, rc;
    memcpy(guess, s->state, sizeof(double) * (size_t)D);
    rc = bdf_residual(s, lower, t, guess, F);
    if (rc) return rc;
    /* jac_finite_diff, bdf.rs:390-411: increment = dt, guess mutated in place */
    const double denom = 1.0 / (s->two * s->dt);
    for (int c = 0; c < D; ++c) {
        guess[c] += s->dt;
        rc = bdf_residual(s, lower, t, guess, above);
        if (rc) return rc;
        guess[c] -= s->two * s->dt;
        rc = bdf_residual(s, lower, t, guess, below);
        if (rc) return rc;
        guess[c] += s->dt;
        for (int r = 0; r < D; ++r)
            AT(s->jac, D, r, c) = (above[r] + s->fd_sign * below[r]) * denom;
    }
    s->n_lu++;
    if (or_inverse(D, s->jac, s->jinv, NULL) != OR_OK) return OR_SINGULAR_MATRIX;
    neg_gemv(D, s->jinv, F, shift);
    for (int d = 0; d < D; ++d) guess[d] += shift[d];
    while (n < 1000) {
        memcpy(Fl, F, sizeof(double) * (size_t)D);
        rc = bdf_residual(s, lower, t, guess, F);
        if (rc) return rc;
        s->n_newton++;
        double diff[MAXD];
        for (int d = 0; d < D; ++d) diff[d] = F[d] - Fl[d];
        neg_gemv(D, s->jinv, diff, adj);
        /* p = (-s^T) adj ; u = s^T Jinv (gemv orders) */
        double pp = adj[0] * (-shift[0]);
        for (int k = 1; k < D; ++k) pp = adj[k] * (-shift[k]) + pp;
        for (int j = 0; j < D; ++j) {
            double acc = AT(s->jinv, D, 0, j) * shift[0];
            for (int k = 1; k < D; ++k) acc = AT(s->jinv, D, k, j) * shift[k] + acc;
            u[j] = acc;
        }
        for (int j = 0; j < D; ++j)
            for (int i = 0; i < D; ++i)
                AT(s->jinv, D, i, j) += (u[j] * (shift[i] + adj[i])) / pp;
        neg_gemv(D, s->jinv, F, shift);
        for (int d = 0; d < D; ++d) guess[d] += shift[d];
        if (nal_norm(D, shift) <= s->tol) return OR_OK;
        n += 1;
    }
    return OR_MAXIMUM_ITERATIONS_EXCEEDED;
}

static int bdf_init(bdf_solver *s, const or_problem *p) {
    memset(s, 0, offsetof(bdf_solver, jinv));
    s->p = p;
    /* Complex state is left out of BDF: the Broyden update (bdf.rs:461-466)
     * multiplies by shift.transpose(), not its conjugate transpose, so a
     * complex system does not run as its real-pair system there. */
    if (or_rhs_is_complex(p->rhs)) return OR_INVALID_ARGUMENT;
    s->O = or_bdf_coefficients(p->method, s->higher, s->lower);
    if (s->O < 0) return OR_INVALID_ARGUMENT;
    s->D = p->dim;
    s->dt_max = p->dt_max;
    s->dt_min = p->dt_min;
    s->time = p->t0;
    s->end = p->t_end;
    s->tol = p->tol;
    s->two = 2.0;
    s->half = 1.0 / s->two;
    s->one_sixth = 1.0 / 6.0;
    s->one_tenth = 1.0 / 10.0;
    s->order = (double)s->O;
    s->dt = (p->dt_max + p->dt_min) * s->half; /* bdf.rs:302 */
    memcpy(s->state, p->y0, sizeof(double) * (size_t)p->dim);
    const int intended = p->semantics == OR_INTENDED;
    s->fd_sign = intended ? -1.0 : 1.0;
    s->lower_own_hist = intended;
    s->f_time_theta = intended ? 1.0 : 0.0;
    s->restore_mul = intended;
    return OR_OK;
}

/* One BDFSolver::step.  On OR_STEP_YIELD, (*yt, yv) is the yielded value. */
static int bdf_step(bdf_solver *s, double *yt, double *yv) {
    const int O = s->O, D = s->D;
    double hi[MAXD], lo[MAXD];
    if (s->yield_memory > 0 && s->yield_memory <= O) { /* bdf.rs:500-512 */
        const int get_item = O - s->yield_memory;
        s->yield_memory -= 1;
        if (s->yield_memory == 0) s->yield_memory = O + 2;
        *yt = hist_t(s, get_item);
        memcpy(yv, hist_at(s, get_item), sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->yield_memory == O + 2) { /* bdf.rs:519-525 */
        s->yield_memory = 0;
        hist_push(s, s->time, s->state);
        hist_pop_front(s);
        *yt = s->time;
        memcpy(yv, s->state, sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->time >= s->end) return OR_STEP_DONE;
    if (s->time + s->dt >= s->end) { /* bdf.rs:531-535 */
        s->dt = s->end - s->time;
        int rc = bdf_runge_kutta(s, 1);
        if (rc) return rc;
        *yt = s->time;
        memcpy(yv, hist_at(s, s->hlen - 1), sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->hlen == 0) { /* bdf.rs:537-546 */
        memcpy(s->save_state, s->state, sizeof(double) * (size_t)D);
        if (s->time + s->dt * s->order >= s->end) s->dt = (s->end - s->time) / s->order;
        int rc = bdf_runge_kutta(s, O);
        if (rc) return rc;
        s->yield_memory = O + 1;
        return OR_STEP_REDO;
    }
    int rc = bdf_secant(s, 0, hi);
    if (rc) return rc;
    rc = bdf_secant(s, 1, lo);
    if (rc) return rc;
    double diff[MAXD];
    for (int d = 0; d < D; ++d) diff[d] = hi[d] - lo[d];
    const double error = nal_norm(D, diff);
    if (error <= s->tol) { /* bdf.rs:583-614 */
        memcpy(s->state, hi, sizeof(double) * (size_t)D);
        s->time += s->dt;
        if (s->yield_memory == O + 1) {
            s->yield_memory -= 1;
            return OR_STEP_REDO;
        }
        hist_push(s, s->time, s->state);
        hist_pop_front(s);
        if (error < s->one_tenth * s->tol) {
            s->dt *= s->two;
            if (s->dt > s->dt_max) s->dt = s->dt_max;
            s->hlen = 0;
            s->hhead = 0;
        }
        *yt = s->time;
        memcpy(yv, s->state, sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    s->n_reject++;
    if (s->yield_memory == O + 1) { /* bdf.rs:620-624 */
        if (s->restore_mul)
            s->time -= s->dt * s->order;
        else
            s->time -= s->dt - s->order; /* Q6 */
        memcpy(s->state, s->save_state, sizeof(double) * (size_t)D);
    }
    s->dt *= s->half;
    if (s->dt < s->dt_min) return OR_MINIMUM_TIME_DELTA_EXCEEDED;
    s->hlen = 0;
    s->hhead = 0;
    return OR_STEP_REDO;
}

static int bdf_solve(const or_problem *p, double *y_out, or_result *r, or_path *path) {
    bdf_solver *s = (bdf_solver *)malloc(sizeof(bdf_solver));
    if (!s) return OR_INVALID_ARGUMENT;
    int rc = bdf_init(s, p);
    if (rc != OR_OK) {
        free(s);
        return rc;
    }
    const int64_t cap = p->max_attempts > 0 ? p->max_attempts : DEFAULT_MAX_ATTEMPTS;
    int64_t calls = 0;
    int status = OR_OK;
    double yt = 0.0, yv[MAXD];
    for (;;) {
        if (calls >= cap) {
            status = OR_MAXIMUM_ITERATIONS_EXCEEDED;
            break;
        }
        calls++;
        const int st = bdf_step(s, &yt, yv);
        if (st == OR_STEP_DONE) break;
        if (st == OR_STEP_YIELD) {
            r->n_yield++;
            r->n_accepted++;
            path_push(path, s->D, yt, yv, 0.0);
        } else if (st == OR_STEP_REDO) {
            /* rejected BDF attempts are counted in bdf_step */
        } else {
            status = st;
            break;
        }
    }
    r->t = s->time;
    r->dt = s->dt;
    r->status = status;
    r->n_rhs = s->n_rhs;
    r->n_newton = s->n_newton;
    r->n_lu = s->n_lu;
    r->n_rejected = s->n_reject;
    memcpy(y_out, s->state, sizeof(double) * (size_t)s->D);
    free(s);
    return status;
}

/* ------------------------------------------------------------------ */
/* Adams-Bashforth-Moulton predictor-corrector (adams.rs:336-633)     */
/* ------------------------------------------------------------------ */

/* x^(1/n) for n = 3 or 5 from exact operations only (the DEVICE profile's
 * replacement for powf(1/order), adams.rs:516,544): x = m 2^e with m in
 * [1, 2); m^(1/n) by a degree-6 polynomial in (m - 1.5) interpolating at
 * Chebyshev nodes (relative error < 2.5e-7); 2^(e/n) = 2^floor(e/n) times the
 * rounded constant 2^((e mod n)/n).  0 -> 0, inf -> inf, NaN -> NaN. */
static double dev_rootn(double x, int n) {
    static const double C3[7] = {1.1447142425533323, 0.254383453864114, -0.05653050145457822,
                                 0.02085714929540036, -0.009260767947395939,
                                 0.0051666994182390956, -0.0027033338703399646};
    static const double C5[7] = {1.0844717711976986, 0.14459836703782772,
                                 -0.038560210159460984, 0.015356114792636277,
                                 -0.007159250830014004, 0.004170036250587015,
                                 -0.002242595967870246};
    static const double S3[3] = {1.0, 1.2599210498948732, 1.5874010519681994};
    static const double S5[5] = {1.0, 1.148698354997035, 1.3195079107728942, 1.515716566510398,
                                 1.7411011265922482};
    if (!(x == x) || x == 0.0 || x == INFINITY) return x;
    const double *c = n == 3 ? C3 : C5;
    const double *sc = n == 3 ? S3 : S5;
    int ex = 0;
    const double m = frexp(x, &ex);
    const double u = fma(m, 2.0, -1.5);
    const int e = ex - 1;
    double p = c[6];
    for (int k = 5; k >= 0; --k) p = fma(p, u, c[k]);
    int a = e / n, b = e % n;
    if (b < 0) {
        b += n;
        a -= 1;
    }
    return ldexp(p * sc[b], a);
}

#define ACAP 16
typedef struct {
    const or_problem *p;
    int O, D, prof, intended;
    double pred[5], corr[5], err_coef;
    double time, dt, end, tol, dt_min, dt_max;
    double two, half, one_sixth, one_tenth, four, order;
    double state[MAXD], save_state[MAXD], implicit[MAXD];
    /* prev_values (t, y) and prev_derivatives deques */
    double vt[ACAP], vy[ACAP][MAXD];
    int vhead, vlen;
    double dv[ACAP][MAXD];
    int dhead, dlen;
    int yield_memory;
    int64_t n_rhs, n_reject;
} adams_solver;

static double *adams_v(adams_solver *s, int k) { return s->vy[(s->vhead + k) % ACAP]; }
static double adams_vt(adams_solver *s, int k) { return s->vt[(s->vhead + k) % ACAP]; }
static double *adams_d(adams_solver *s, int k) { return s->dv[(s->dhead + k) % ACAP]; }
static void adams_push_v(adams_solver *s, double t, const double *y) {
    const int slot = (s->vhead + s->vlen) % ACAP;
    s->vt[slot] = t;
    memcpy(s->vy[slot], y, sizeof(double) * (size_t)s->D);
    s->vlen++;
}
static void adams_push_d(adams_solver *s, const double *f) {
    const int slot = (s->dhead + s->dlen) % ACAP;
    memcpy(s->dv[slot], f, sizeof(double) * (size_t)s->D);
    s->dlen++;
}
static void adams_pop_v(adams_solver *s) {
    if (s->vlen) {
        s->vhead = (s->vhead + 1) % ACAP;
        s->vlen--;
    }
}
static void adams_pop_d(adams_solver *s) {
    if (s->dlen) {
        s->dhead = (s->dhead + 1) % ACAP;
        s->dlen--;
    }
}

static int adams_rhs(adams_solver *s, double t, const double *y, double *f) {
    s->n_rhs++;
    int rc = or_rhs_eval(s->p->rhs, s->prof, s->D, t, y, s->p->params, s->p->shared, f);
    if (rc == OR_INVALID_ARGUMENT) return rc;
    return rc == OR_OK ? OR_OK : OR_USER_ERROR;
}

/* adams.rs:344-396: classical RK4; states (and their derivatives) after
 * steps 1..n-1 and the final one are pushed onto the deques */
static int adams_runge_kutta(adams_solver *s, int iterations) {
    const int D = s->D, pr = s->prof;
    double k1[MAXD], k2[MAXD], k3[MAXD], k4[MAXD], mid[MAXD], f[MAXD];
    for (int it = 0; it < iterations; ++it) {
        int rc = adams_rhs(s, s->time, s->state, k1);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k1[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = madd(pr, k1[d], s->half, s->state[d]);
        rc = adams_rhs(s, madd(pr, s->half, s->dt, s->time), mid, k2);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k2[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = madd(pr, k2[d], s->half, s->state[d]);
        rc = adams_rhs(s, madd(pr, s->half, s->dt, s->time), mid, k3);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k3[d] *= s->dt;
        for (int d = 0; d < D; ++d) mid[d] = s->state[d] + k3[d];
        rc = adams_rhs(s, s->time + s->dt, mid, k4);
        if (rc) return rc;
        for (int d = 0; d < D; ++d) k4[d] *= s->dt;
        if (it != 0) {
            rc = adams_rhs(s, s->time, s->state, f);
            if (rc) return rc;
            adams_push_d(s, f);
            adams_push_v(s, s->time, s->state);
        }
        for (int d = 0; d < D; ++d) {
            double sum = madd(pr, k2[d], s->two, k1[d]);
            sum = madd(pr, k3[d], s->two, sum);
            sum = sum + k4[d];
            s->state[d] = madd(pr, sum, s->one_sixth, s->state[d]);
        }
        s->time += s->dt;
    }
    int rc = adams_rhs(s, s->time, s->state, f);
    if (rc) return rc;
    adams_push_d(s, f);
    adams_push_v(s, s->time, s->state);
    return OR_OK;
}

static int adams_init(adams_solver *s, const or_problem *p) {
    memset(s, 0, sizeof *s);
    s->p = p;
    s->O = p->method == OR_ADAMS5 ? 5 : 3;
    s->D = p->dim;
    s->prof = p->profile;
    s->intended = p->semantics == OR_INTENDED;
    if (s->O == 5) { /* adams.rs:573-600 */
        const double tf = 24.0, sh = 720.0;
        s->pred[0] = 55.0 / tf; s->pred[1] = -59.0 / tf; s->pred[2] = 37.0 / tf;
        s->pred[3] = -9.0 / tf; s->pred[4] = 0.0;
        s->corr[0] = 251.0 / sh; s->corr[1] = 646.0 / sh; s->corr[2] = -264.0 / sh;
        s->corr[3] = 106.0 / sh; s->corr[4] = -19.0 / sh;
    } else { /* adams.rs:640-661 */
        s->pred[0] = 1.0 + 1.0 / 2.0; s->pred[1] = -(1.0 / 2.0); s->pred[2] = 0.0;
        s->corr[0] = 5.0 / 12.0; s->corr[1] = 2.0 / 3.0; s->corr[2] = -(1.0 / 12.0);
    }
    s->err_coef = 19.0 / 270.0;
    s->two = 2.0;
    s->half = 1.0 / s->two;
    s->one_sixth = 1.0 / 6.0;
    s->one_tenth = 1.0 / 10.0;
    s->four = s->two * s->two;
    s->order = (double)s->O;
    s->dt_max = p->dt_max;
    s->dt_min = p->dt_min;
    s->time = p->t0;
    s->end = p->t_end;
    s->tol = p->tol;
    s->dt = (p->dt_max + p->dt_min) * s->half; /* adams.rs:288 */
    memcpy(s->state, p->y0, sizeof(double) * (size_t)p->dim);
    return OR_OK;
}

/* q = (tol / (2 error))^(1/order) (adams.rs:516, 544) */
static double adams_q(adams_solver *s, double error) {
    const double r = s->tol / (s->two * error);
    return s->prof == OR_PROFILE_DEVICE ? dev_rootn(r, s->O) : pow(r, 1.0 / s->order);
}

/* One AdamsSolver::step (adams.rs:411-561). */
static int adams_step(adams_solver *s, double *yt, double *yv) {
    const int O = s->O, D = s->D, pr = s->prof;
    if (s->yield_memory > 0 && s->yield_memory < O) { /* adams.rs:415-428 */
        const int item = O - s->yield_memory - 1;
        s->yield_memory -= 1;
        if (s->yield_memory == 0) s->yield_memory = O + 1;
        *yt = adams_vt(s, item);
        memcpy(yv, adams_v(s, item), sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->yield_memory == O + 1) { /* adams.rs:434-441 */
        s->yield_memory = 0;
        adams_push_v(s, s->time, s->state);
        adams_pop_v(s);
        if (s->intended) {
            /* the step's derivative joins the history too; as built the
             * derivative deque keeps the bootstrap's (one step stale) */
            adams_push_d(s, s->implicit);
            adams_pop_d(s);
        }
        *yt = s->time;
        memcpy(yv, s->state, sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->time >= s->end) return OR_STEP_DONE;
    if (s->time + s->dt >= s->end) { /* adams.rs:447-451 */
        s->dt = s->end - s->time;
        const int rc = adams_runge_kutta(s, 1);
        if (rc) return rc;
        *yt = s->time;
        memcpy(yv, adams_v(s, s->vlen - 1), sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    if (s->vlen == 0) { /* adams.rs:453-464 */
        memcpy(s->save_state, s->state, sizeof(double) * (size_t)D);
        if (s->time + s->dt * (s->order - 1.0) >= s->end)
            s->dt = (s->end - s->time) / (s->order - 1.0);
        const int rc = adams_runge_kutta(s, O - 1);
        if (rc) return rc;
        s->yield_memory = O;
        return OR_STEP_REDO;
    }
    /* adams.rs:466-486 */
    double scratch[MAXD], predictor[MAXD], corrector[MAXD], diff[MAXD];
    for (int d = 0; d < D; ++d) scratch[d] = adams_d(s, 0)[d] * s->pred[O - 2];
    for (int i = 1; i < O - 1; ++i) {
        const double c = s->pred[O - i - 2];
        for (int d = 0; d < D; ++d) scratch[d] = madd(pr, adams_d(s, i)[d], c, scratch[d]);
    }
    for (int d = 0; d < D; ++d) predictor[d] = madd(pr, scratch[d], s->dt, s->state[d]);
    int rc = adams_rhs(s, s->time + s->dt, predictor, s->implicit);
    if (rc) return rc;
    for (int d = 0; d < D; ++d) scratch[d] = s->implicit[d] * s->corr[0];
    for (int i = 0; i < O - 1; ++i) {
        const double c = s->corr[O - i - 1];
        for (int d = 0; d < D; ++d) scratch[d] = madd(pr, adams_d(s, i)[d], c, scratch[d]);
    }
    for (int d = 0; d < D; ++d) corrector[d] = madd(pr, scratch[d], s->dt, s->state[d]);
    for (int d = 0; d < D; ++d) diff[d] = corrector[d] - predictor[d];
    const double nrm = pr == OR_PROFILE_DEVICE  ? sqrt(dev_sumsq(D, diff))
                       : or_rhs_is_complex(s->p->rhs) ? nal_norm_c(D / 2, diff)
                                                      : nal_norm(D, diff);
    const double error = s->err_coef / s->dt * nrm;
    if (error <= s->tol) { /* adams.rs:490-525 */
        memcpy(s->state, corrector, sizeof(double) * (size_t)D);
        s->time += s->dt;
        if (s->yield_memory == O) {
            s->yield_memory -= 1;
            return OR_STEP_REDO;
        }
        adams_push_d(s, s->implicit);
        adams_push_v(s, s->time, s->state);
        adams_pop_v(s);
        adams_pop_d(s);
        if (error < s->one_tenth * s->tol) {
            const double q = adams_q(s, error);
            if (q > s->four)
                s->dt *= s->four;
            else
                s->dt *= q;
            if (s->dt > s->dt_max) s->dt = s->dt_max;
            s->vlen = s->vhead = 0;
            s->dlen = s->dhead = 0;
        }
        *yt = s->time;
        memcpy(yv, s->state, sizeof(double) * (size_t)D);
        return OR_STEP_YIELD;
    }
    s->n_reject++;
    if (s->yield_memory == O) { /* adams.rs:531-536 */
        s->time -= s->dt * (s->order - 1.0);
        memcpy(s->state, s->save_state, sizeof(double) * (size_t)D);
    }
    const double q = adams_q(s, error); /* adams.rs:538-556 */
    if (q < s->one_tenth)
        s->dt *= s->one_tenth;
    else
        s->dt *= q;
    if (s->dt < s->dt_min) return OR_MINIMUM_TIME_DELTA_EXCEEDED;
    s->vlen = s->vhead = 0;
    s->dlen = s->dhead = 0;
    return OR_STEP_REDO;
}

int32_t or_adams_coefficients(int32_t method, double *pred, double *corr, double *err) {
    if (method != OR_ADAMS5 && method != OR_ADAMS3) return OR_INVALID_ARGUMENT;
    or_problem p;
    memset(&p, 0, sizeof p);
    p.method = method;
    p.dim = 1;
    adams_solver *s = (adams_solver *)malloc(sizeof(adams_solver));
    if (!s) return OR_INVALID_ARGUMENT;
    double y0 = 0.0;
    p.y0 = &y0;
    adams_init(s, &p);
    const int O = s->O;
    memcpy(pred, s->pred, sizeof(double) * (size_t)O);
    memcpy(corr, s->corr, sizeof(double) * (size_t)O);
    *err = s->err_coef;
    free(s);
    return O;
}

double or_rootn(double x, int32_t n) { return dev_rootn(x, n); }
double or_root8(double r) { return dev_root8(r); }

static int adams_solve(const or_problem *p, double *y_out, or_result *r, or_path *path) {
    adams_solver *s = (adams_solver *)malloc(sizeof(adams_solver));
    if (!s) return OR_INVALID_ARGUMENT;
    int rc = adams_init(s, p);
    if (rc != OR_OK) {
        free(s);
        return rc;
    }
    const int64_t cap = p->max_attempts > 0 ? p->max_attempts : DEFAULT_MAX_ATTEMPTS;
    int64_t calls = 0;
    int status = OR_OK;
    double yt = 0.0, yv[MAXD];
    for (;;) {
        if (calls >= cap) {
            status = OR_MAXIMUM_ITERATIONS_EXCEEDED;
            break;
        }
        calls++;
        const int st = adams_step(s, &yt, yv);
        if (st == OR_STEP_DONE) break;
        if (st == OR_STEP_YIELD) {
            r->n_yield++;
            r->n_accepted++;
            path_push(path, s->D, yt, yv, 0.0);
        } else if (st != OR_STEP_REDO) {
            status = st;
            break;
        }
    }
    r->t = s->time;
    r->dt = s->dt;
    r->status = status;
    r->n_rhs = s->n_rhs;
    r->n_rejected = s->n_reject;
    memcpy(y_out, s->state, sizeof(double) * (size_t)s->D);
    free(s);
    return status;
}

/* ------------------------------------------------------------------ */
/* public entry points                                                 */
/* ------------------------------------------------------------------ */

int32_t or_solve(const or_problem *p, double *y_out, or_result *r, or_path *path,
                 double *stop_y) {
    if (!p || !y_out || !r || !p->y0 || p->dim < 1 || p->dim > MAXD)
        return OR_INVALID_ARGUMENT;
    memset(r, 0, sizeof *r);
    if (path) path->len = 0;
    if (p->method == OR_RK45 || p->method == OR_RK23)
        return rk_solve(p, y_out, r, path, stop_y);
    if (p->method == OR_BDF6 || p->method == OR_BDF2) {
        if (p->n_stops) return OR_INVALID_ARGUMENT;
        return bdf_solve(p, y_out, r, path);
    }
    if (p->method == OR_ADAMS5 || p->method == OR_ADAMS3) {
        if (p->n_stops) return OR_INVALID_ARGUMENT;
        return adams_solve(p, y_out, r, path);
    }
    return OR_INVALID_ARGUMENT;
}

int32_t or_solve_ensemble(const or_problem *tmpl, int64_t n, int32_t n_params,
                          const double *y0, const double *params,
                          const double *t_end, const double *tol, double *t_out,
                          double *y_out, int32_t *status, int64_t *n_accepted,
                          int64_t *n_rejected, int32_t threads) {
    if (!tmpl || n < 0 || tmpl->dim < 1 || tmpl->dim > MAXD) return OR_INVALID_ARGUMENT;
    const int D = tmpl->dim;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < n; ++i) {
        or_problem p = *tmpl;
        double yy[MAXD], pr[64], yo[MAXD];
        for (int d = 0; d < D; ++d) yy[d] = y0[(size_t)d * n + i];
        for (int k = 0; k < n_params && k < 64; ++k) pr[k] = params[(size_t)k * n + i];
        p.y0 = yy;
        p.params = n_params > 0 ? pr : tmpl->params;
        if (t_end) p.t_end = t_end[i];
        if (tol) p.tol = tol[i];
        or_result r;
        or_solve(&p, yo, &r, NULL, NULL);
        if (t_out) t_out[i] = r.t;
        if (y_out)
            for (int d = 0; d < D; ++d) y_out[(size_t)d * n + i] = yo[d];
        if (status) status[i] = r.status;
        if (n_accepted) n_accepted[i] = r.n_accepted;
        if (n_rejected) n_rejected[i] = r.n_rejected;
    }
    (void)threads;
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* builder restatement (rk.rs:136-256; bdf.rs:144-264 is identical)    */
/* ------------------------------------------------------------------ */

int32_t or_builder_new(or_builder *b, int32_t method, int32_t static_dim) {
    memset(b, 0, sizeof *b);
    b->method = method;
    /* Dimension::dim() on Dyn fails (lib.rs:70-72); static_dim == 0 models Dyn */
    if (static_dim <= 0) return OR_STATIC_ON_DYNAMIC;
    b->dim = static_dim;
    b->built = 1;
    return OR_OK;
}

int32_t or_builder_new_dyn(or_builder *b, int32_t method, int32_t size) {
    memset(b, 0, sizeof *b);
    b->method = method;
    b->is_dyn = 1;
    b->dim = size;
    b->built = 1;
    return OR_OK;
}

int32_t or_builder_with_tolerance(or_builder *b, double tol) {
    if (tol <= 0.0) return OR_TOLERANCE_OOB; /* rk.rs:169 (NaN passes) */
    b->tol = tol;
    b->have_tol = 1;
    return OR_OK;
}

int32_t or_builder_with_maximum_dt(or_builder *b, double max) {
    if (max <= 0.0) return OR_TIME_DELTA_OOB;
    b->dt_max = max;
    b->have_dt_max = 1;
    if (b->have_dt_min && b->dt_min > max) b->dt_min = max;
    return OR_OK;
}

int32_t or_builder_with_minimum_dt(or_builder *b, double min) {
    if (min <= 0.0) return OR_TIME_DELTA_OOB;
    b->dt_min = min;
    b->have_dt_min = 1;
    if (b->have_dt_max && b->dt_max < min) b->dt_max = min;
    return OR_OK;
}

int32_t or_builder_with_initial_time(or_builder *b, double v) {
    b->t0 = v; /* stored before the check (rk.rs:213) */
    b->have_t0 = 1;
    if (b->have_t_end && b->t_end <= v) return OR_TIME_START_OOB;
    return OR_OK;
}

int32_t or_builder_with_ending_time(or_builder *b, double v) {
    b->t_end = v;
    b->have_t_end = 1;
    if (b->have_t0 && b->t0 >= v) return OR_TIME_END_OOB;
    return OR_OK;
}

int32_t or_builder_with_initial_conditions(or_builder *b, int32_t n) {
    if (n != b->dim) return OR_INVALID_ARGUMENT; /* nalgebra panics here */
    b->have_y0 = 1;
    return OR_OK;
}

int32_t or_builder_with_derivative(or_builder *b) {
    b->have_deriv = 1;
    return OR_OK;
}

int32_t or_builder_solve(const or_builder *b) {
    /* rk.rs:250-256: the first missing parameter, in this order */
    if (!b->have_dt_max || !b->have_dt_min || !b->have_tol || !b->have_t0 ||
        !b->have_t_end || !b->have_y0 || !b->have_deriv)
        return OR_MISSING_PARAMETERS;
    return OR_OK;
}
