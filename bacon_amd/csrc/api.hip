// api.hip — the extern "C" boundary of libbacon_gpu.so.
//
// Builder semantics follow bacon-sci's IVPSolver implementations
// (src/ivp/rk.rs:136-256, identical in src/ivp/bdf.rs:144-264): eager
// validation in each setter, MissingParameters checked by solve() in the
// order dt_max, dt_min, tolerance, start, end, state, derivative.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "bg_internal.h"

namespace bg {

static std::mutex g_dev_mutex;
static int g_cu_count[64];

int device_cu_count(int dev) {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if (dev < 0 || dev >= 64) return 1;
    if (!g_cu_count[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n < 1)
            n = 1;
        g_cu_count[dev] = n;
    }
    return g_cu_count[dev];
}

static bool is_rk(int m) { return m == BG_RK45 || m == BG_RK23; }
static bool is_bdf(int m) { return m == BG_BDF6 || m == BG_BDF2; }
static bool is_adams(int m) { return m == BG_ADAMS5 || m == BG_ADAMS3; }
static bool valid_method(int m) { return is_rk(m) || is_bdf(m) || is_adams(m); }

// number of per-instance parameters a registered RHS reads
static int rhs_param_count(int rhs) {
    switch (rhs) {
    case BG_RHS_LORENZ: return 3;
    case BG_RHS_BRUSSELATOR: return 3;
    case BG_RHS_ROBERTSON: return 3;
    case BG_RHS_TWO_LEVEL: return 3;
    case BG_RHS_LOTKA_VOLTERRA: return 4;
    case BG_RHS_FAIL_AFTER: return 1;
    case BG_RHS_EXP: case BG_RHS_NEG: case BG_RHS_QUADRATIC: case BG_RHS_COS:
    case BG_RHS_OSCILLATOR: case BG_RHS_LINEAR: return 0;
    default: return -1;
    }
}

}  // namespace bg

using namespace bg;

// ===========================================================================
// single-instance builder + iterator
// ===========================================================================
struct bg_solver {
    int method, dim_kind;
    uint32_t dim;
    int semantics = BG_INTENDED;
    bool have_tol = false, have_dt_max = false, have_dt_min = false, have_t0 = false,
         have_t_end = false, have_y0 = false, have_deriv = false;
    double tol = 0, dt_max = 0, dt_min = 0, t0 = 0, t_end = 0;
    std::vector<double> y0, params;
    int rhs = -1;
    // iterator
    bool solved = false, fused = false;
    int failure = BG_OK;
    std::vector<double> path_t, path_y;
    size_t next = 0;
    double last_t = 0;
};

extern "C" int bg_new(int method, int dim_kind, uint32_t dim, bg_solver **out) {
    if (!out || !valid_method(method)) return BG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (dim_kind == BG_DIM_DYN) return BG_STATIC_ON_DYNAMIC;  // lib.rs:70-72
    if (dim_kind != BG_DIM_STATIC || dim == 0 || dim > 128) return BG_ERR_INVALID_ARGUMENT;
    bg_solver *s = new bg_solver();
    s->method = method;
    s->dim_kind = dim_kind;
    s->dim = dim;
    *out = s;
    return BG_OK;
}

extern "C" int bg_new_dyn(int method, int dim_kind, uint32_t size, bg_solver **out) {
    if (!out || !valid_method(method)) return BG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (dim_kind == BG_DIM_STATIC) return BG_DYNAMIC_ON_STATIC;  // lib.rs:62-64
    if (dim_kind != BG_DIM_DYN || size == 0 || size > 128) return BG_ERR_INVALID_ARGUMENT;
    bg_solver *s = new bg_solver();
    s->method = method;
    s->dim_kind = dim_kind;
    s->dim = size;
    *out = s;
    return BG_OK;
}

extern "C" int bg_dim(const bg_solver *s, uint32_t *dim) {
    if (!s || !dim) return BG_ERR_INVALID_ARGUMENT;
    *dim = s->dim;
    return BG_OK;
}

extern "C" int bg_with_tolerance(bg_solver *s, double tol) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    if (tol <= 0.0) return BG_TOLERANCE_OOB;  // rk.rs:168-174
    s->tol = tol;
    s->have_tol = true;
    return BG_OK;
}

extern "C" int bg_with_maximum_dt(bg_solver *s, double max) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    if (max <= 0.0) return BG_TIME_DELTA_OOB;  // rk.rs:179-192
    s->dt_max = max;
    s->have_dt_max = true;
    if (s->have_dt_min && s->dt_min > max) s->dt_min = max;
    return BG_OK;
}

extern "C" int bg_with_minimum_dt(bg_solver *s, double min) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    if (min <= 0.0) return BG_TIME_DELTA_OOB;  // rk.rs:197-210
    s->dt_min = min;
    s->have_dt_min = true;
    if (s->have_dt_max && s->dt_max < min) s->dt_max = min;
    return BG_OK;
}

extern "C" int bg_with_initial_time(bg_solver *s, double initial) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    s->t0 = initial;  // stored before the check (rk.rs:212-222)
    s->have_t0 = true;
    if (s->have_t_end && s->t_end <= initial) return BG_TIME_START_OOB;
    return BG_OK;
}

extern "C" int bg_with_ending_time(bg_solver *s, double ending) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    s->t_end = ending;  // rk.rs:224-234
    s->have_t_end = true;
    if (s->have_t0 && s->t0 >= ending) return BG_TIME_END_OOB;
    return BG_OK;
}

extern "C" int bg_with_initial_conditions_slice(bg_solver *s, const double *y0, size_t n) {
    if (!s || (!y0 && n)) return BG_ERR_INVALID_ARGUMENT;
    // the reference panics in from_column_slice_generic on a length mismatch
    if (n != s->dim) return BG_ERR_INVALID_ARGUMENT;
    s->y0.assign(y0, y0 + n);
    s->have_y0 = true;
    return BG_OK;
}

extern "C" int bg_with_derivative(bg_solver *s, int rhs, const double *params, size_t n_params) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    const int P = rhs_param_count(rhs);
    if (P < 0) return BG_ERR_UNSUPPORTED;
    if ((size_t)P > n_params || (n_params && !params) || n_params > 8)
        return BG_ERR_INVALID_ARGUMENT;
    s->rhs = rhs;
    s->params.assign(params, params + n_params);
    s->have_deriv = true;
    return BG_OK;
}

extern "C" int bg_with_semantics(bg_solver *s, int semantics) {
    if (!s || (semantics != BG_AS_BUILT && semantics != BG_INTENDED))
        return BG_ERR_INVALID_ARGUMENT;
    s->semantics = semantics;
    return BG_OK;
}

extern "C" int bg_with_dt_min(bg_solver *s, double v) { return bg_with_minimum_dt(s, v); }
extern "C" int bg_with_dt_max(bg_solver *s, double v) { return bg_with_maximum_dt(s, v); }
extern "C" int bg_with_start(bg_solver *s, double v) { return bg_with_initial_time(s, v); }
extern "C" int bg_with_end(bg_solver *s, double v) { return bg_with_ending_time(s, v); }

// Drive a resumable single-instance path kernel: launch, drain the path
// buffer into the solver's host path, relaunch until the state says finished.
template <class State, class Launch>
static int drain_path(bg_solver *s, const State &init, Launch &&launch) {
    const uint32_t cap = 1u << 16;
    const uint32_t D = s->dim;
    State host = init;
    State *dstate = nullptr;
    double *dpt = nullptr, *dpy = nullptr;
    int rc = BG_OK;
    if (hipMalloc((void **)&dstate, sizeof(State)) != hipSuccess ||
        hipMalloc((void **)&dpt, sizeof(double) * cap) != hipSuccess ||
        hipMalloc((void **)&dpy, sizeof(double) * cap * D) != hipSuccess)
        rc = BG_ERR_HIP;
    if (rc == BG_OK && hipMemcpy(dstate, &host, sizeof host, hipMemcpyHostToDevice) != hipSuccess)
        rc = BG_ERR_HIP;
    s->path_t.clear();
    s->path_y.clear();
    while (rc == BG_OK) {
        rc = launch(dstate, dpt, dpy, cap);
        if (rc != BG_OK) break;
        if (hipMemcpy(&host, dstate, sizeof host, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
        const size_t old = s->path_t.size();
        s->path_t.resize(old + host.count);
        s->path_y.resize((old + host.count) * D);
        if (host.count &&
            (hipMemcpy(s->path_t.data() + old, dpt, sizeof(double) * host.count,
                       hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(s->path_y.data() + old * D, dpy, sizeof(double) * host.count * D,
                       hipMemcpyDeviceToHost) != hipSuccess)) {
            rc = BG_ERR_HIP;
            break;
        }
        if (host.finished) break;
    }
    if (dstate) hipFree(dstate);
    if (dpt) hipFree(dpt);
    if (dpy) hipFree(dpy);
    if (rc == BG_OK) s->failure = host.status;
    return rc;
}

static int solve_rk_path(bg_solver *s) {
    if (!rk_supported(s->method, s->rhs, s->dim)) return BG_ERR_UNSUPPORTED;
    PathState host;
    memset(&host, 0, sizeof host);
    host.t = s->t0;
    host.end = s->t_end;
    host.tol = s->tol;
    host.dt_min = s->dt_min;
    host.dt_max = s->dt_max;
    host.dt = (s->dt_max + s->dt_min) * 0.5;  // rk.rs:315
    host.max_attempts = 10000000ull;
    for (uint32_t k = 0; k < s->dim; ++k) host.y[k] = s->y0[k];
    for (size_t k = 0; k < s->params.size() && k < 8; ++k) host.p[k] = s->params[k];
    return drain_path(s, host, [&](PathState *st, double *pt, double *py, uint32_t cap) {
        return rk_launch_path(s->method, s->rhs, s->semantics, s->dim, st, pt, py, cap, nullptr);
    });
}

static int solve_bdf_path(bg_solver *s) {
    if (!bdf_supported(s->method, s->rhs, s->dim)) return BG_ERR_UNSUPPORTED;
    BdfPathState host;
    memset(&host, 0, sizeof host);
    host.t = s->t0;
    host.end = s->t_end;
    host.tol = s->tol;
    host.dt_min = s->dt_min;
    host.dt_max = s->dt_max;
    host.dt = (s->dt_max + s->dt_min) * 0.5;  // bdf.rs:302
    host.max_steps = 1000000ull;
    for (uint32_t k = 0; k < s->dim; ++k) host.y[k] = host.save[k] = s->y0[k];
    for (size_t k = 0; k < s->params.size() && k < 8; ++k) host.p[k] = s->params[k];
    return drain_path(s, host, [&](BdfPathState *st, double *pt, double *py, uint32_t cap) {
        return bdf_launch_path(s->method, s->rhs, s->semantics, s->dim, st, pt, py, cap, nullptr);
    });
}

static int solve_adams_path(bg_solver *s) {
    if (!adams_supported(s->method, s->rhs, s->dim)) return BG_ERR_UNSUPPORTED;
    AdamsPathState host;
    memset(&host, 0, sizeof host);
    host.t = s->t0;
    host.end = s->t_end;
    host.tol = s->tol;
    host.dt_min = s->dt_min;
    host.dt_max = s->dt_max;
    host.dt = (s->dt_max + s->dt_min) * 0.5;  // adams.rs:288
    host.max_steps = 10000000ull;
    for (uint32_t k = 0; k < s->dim; ++k) host.y[k] = s->y0[k];
    for (size_t k = 0; k < s->params.size() && k < 8; ++k) host.p[k] = s->params[k];
    return drain_path(s, host, [&](AdamsPathState *st, double *pt, double *py, uint32_t cap) {
        return adams_launch_path(s->method, s->rhs, s->semantics, s->dim, st, pt, py, cap,
                                 nullptr);
    });
}

extern "C" int bg_solve(bg_solver *s) {
    if (!s) return BG_ERR_INVALID_ARGUMENT;
    // rk.rs:250-256 / bdf.rs:258-264 / adams.rs:250-256
    if (!s->have_dt_max || !s->have_dt_min || !s->have_tol || !s->have_t0 || !s->have_t_end ||
        !s->have_y0 || !s->have_deriv)
        return BG_MISSING_PARAMETERS;
    s->solved = false;
    s->fused = false;
    s->next = 0;
    const int rc = is_rk(s->method)      ? solve_rk_path(s)
                   : is_adams(s->method) ? solve_adams_path(s)
                                         : solve_bdf_path(s);
    if (rc != BG_OK) return rc;
    s->solved = true;
    return BG_OK;
}

extern "C" int bg_next(bg_solver *s, double *t, double *y) {
    if (!s || !t || !y) return BG_ERR_INVALID_ARGUMENT;
    if (!s->solved) return BG_ERR_STATE;
    if (s->fused) return BG_DONE;
    if (s->next < s->path_t.size()) {
        *t = s->path_t[s->next];
        memcpy(y, s->path_y.data() + s->next * s->dim, sizeof(double) * s->dim);
        s->last_t = *t;
        s->next++;
        return BG_OK;
    }
    s->fused = true;  // ivp.rs:232-235: the first failure ends the iteration
    return s->failure != BG_OK ? s->failure : BG_DONE;
}

extern "C" int bg_time(const bg_solver *s, double *t) {
    if (!s || !t) return BG_ERR_INVALID_ARGUMENT;
    *t = s->next ? s->last_t : s->t0;
    return BG_OK;
}

extern "C" void bg_destroy(bg_solver *s) { delete s; }

// ===========================================================================
// batched builder
// ===========================================================================
namespace {
struct Field {
    bool set = false;
    double scalar = 0;
    std::vector<double> v;  // empty = broadcast scalar
    double at(size_t i) const { return v.empty() ? scalar : v[i]; }
    void expand(size_t n) {
        if (v.empty()) v.assign(n, scalar);
    }
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes && p) return BG_OK;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        if (b == 0) return BG_OK;
        if (hipMalloc(&p, b) != hipSuccess) return BG_ERR_HIP;
        bytes = b;
        return BG_OK;
    }
    ~DevBuf() {
        if (p) hipFree(p);
    }
};
}  // namespace

struct bg_ensemble {
    int method;
    uint32_t dim;
    uint64_t n;
    int semantics = BG_INTENDED;
    int rhs = -1;
    uint32_t n_params = 0;
    uint64_t max_attempts = 0;
    int lu_bits = 64;
    Field tol, dt_max, dt_min, t0, t_end;
    std::vector<double> y0, params, shared;
    bool have_y0 = false, have_deriv = false, solved = false;
    DevBuf d_y0, d_params, d_shared, d_t0, d_tend, d_tol, d_dtmin, d_dtmax, d_order;
    DevBuf d_tout, d_yout, d_status, d_counts, d_work;
};

static int set_field(bg_ensemble *e, Field &f, const double *v, uint64_t count) {
    if (!e || !v || (count != 1 && count != e->n)) return BG_ERR_INVALID_ARGUMENT;
    if (count == 1) {
        f.scalar = v[0];
        f.v.clear();
    } else {
        f.v.assign(v, v + count);
    }
    f.set = true;
    return BG_OK;
}

extern "C" int bg_ensemble_new(int method, uint32_t dim, uint64_t n, bg_ensemble **out) {
    if (!out || !valid_method(method) || dim == 0 || dim > 128 || n > 0x7fffffffull)
        return BG_ERR_INVALID_ARGUMENT;
    bg_ensemble *e = new bg_ensemble();
    e->method = method;
    e->dim = dim;
    e->n = n;
    *out = e;
    return BG_OK;
}

extern "C" int bg_ensemble_with_tolerance(bg_ensemble *e, const double *tol, uint64_t count) {
    if (!e || !tol || (count != 1 && count != e->n)) return BG_ERR_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < count; ++i)
        if (tol[i] <= 0.0) return BG_TOLERANCE_OOB;
    return set_field(e, e->tol, tol, count);
}

extern "C" int bg_ensemble_with_maximum_dt(bg_ensemble *e, const double *v, uint64_t count) {
    if (!e || !v || (count != 1 && count != e->n)) return BG_ERR_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < count; ++i)
        if (v[i] <= 0.0) return BG_TIME_DELTA_OOB;
    set_field(e, e->dt_max, v, count);
    if (e->dt_min.set) {  // lower a larger minimum (rk.rs:184-189)
        bool any = false;
        for (uint64_t i = 0; i < e->n && !any; ++i) any = e->dt_min.at(i) > e->dt_max.at(i);
        if (any) {
            if (e->dt_min.v.empty() && e->dt_max.v.empty()) {
                e->dt_min.scalar = e->dt_max.scalar;
            } else {
                e->dt_min.expand(e->n);
                for (uint64_t i = 0; i < e->n; ++i)
                    if (e->dt_min.v[i] > e->dt_max.at(i)) e->dt_min.v[i] = e->dt_max.at(i);
            }
        }
    }
    return BG_OK;
}

extern "C" int bg_ensemble_with_minimum_dt(bg_ensemble *e, const double *v, uint64_t count) {
    if (!e || !v || (count != 1 && count != e->n)) return BG_ERR_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < count; ++i)
        if (v[i] <= 0.0) return BG_TIME_DELTA_OOB;
    set_field(e, e->dt_min, v, count);
    if (e->dt_max.set) {  // raise a smaller maximum (rk.rs:202-207)
        bool any = false;
        for (uint64_t i = 0; i < e->n && !any; ++i) any = e->dt_max.at(i) < e->dt_min.at(i);
        if (any) {
            if (e->dt_min.v.empty() && e->dt_max.v.empty()) {
                e->dt_max.scalar = e->dt_min.scalar;
            } else {
                e->dt_max.expand(e->n);
                for (uint64_t i = 0; i < e->n; ++i)
                    if (e->dt_max.v[i] < e->dt_min.at(i)) e->dt_max.v[i] = e->dt_min.at(i);
            }
        }
    }
    return BG_OK;
}

extern "C" int bg_ensemble_with_initial_time(bg_ensemble *e, const double *v, uint64_t count) {
    int rc = set_field(e, e->t0, v, count);
    if (rc != BG_OK) return rc;
    if (e->t_end.set)
        for (uint64_t i = 0; i < e->n; ++i)
            if (e->t_end.at(i) <= e->t0.at(i)) return BG_TIME_START_OOB;
    return BG_OK;
}

extern "C" int bg_ensemble_with_ending_time(bg_ensemble *e, const double *v, uint64_t count) {
    int rc = set_field(e, e->t_end, v, count);
    if (rc != BG_OK) return rc;
    if (e->t0.set)
        for (uint64_t i = 0; i < e->n; ++i)
            if (e->t0.at(i) >= e->t_end.at(i)) return BG_TIME_END_OOB;
    return BG_OK;
}

extern "C" int bg_ensemble_with_initial_conditions(bg_ensemble *e, const double *y0,
                                                   uint64_t count) {
    if (!e || (!y0 && count) || count != (uint64_t)e->dim * e->n) return BG_ERR_INVALID_ARGUMENT;
    e->y0.assign(y0, y0 + count);
    e->have_y0 = true;
    return BG_OK;
}

extern "C" int bg_ensemble_with_derivative(bg_ensemble *e, int rhs, const double *params,
                                           uint32_t n_params, const double *shared,
                                           uint64_t n_shared) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    const int P = rhs_param_count(rhs);
    if (P < 0) return BG_ERR_UNSUPPORTED;
    if (n_params < (uint32_t)P || (n_params && !params) || (n_shared && !shared))
        return BG_ERR_INVALID_ARGUMENT;
    // the linear right-hand side reads A (dim x dim, row-major) then b
    if (rhs == BG_RHS_LINEAR && n_shared < (uint64_t)e->dim * e->dim + e->dim)
        return BG_ERR_INVALID_ARGUMENT;
    e->rhs = rhs;
    e->n_params = n_params;
    e->params.assign(params, params + (size_t)n_params * e->n);
    e->shared.assign(shared, shared + n_shared);
    e->have_deriv = true;
    return BG_OK;
}

extern "C" int bg_ensemble_with_semantics(bg_ensemble *e, int semantics) {
    if (!e || (semantics != BG_AS_BUILT && semantics != BG_INTENDED))
        return BG_ERR_INVALID_ARGUMENT;
    e->semantics = semantics;
    return BG_OK;
}

extern "C" int bg_ensemble_with_max_attempts(bg_ensemble *e, uint64_t max_attempts) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    e->max_attempts = max_attempts;
    return BG_OK;
}

extern "C" int bg_ensemble_with_lu_precision(bg_ensemble *e, int bits) {
    if (!e || (bits != 32 && bits != 64)) return BG_ERR_INVALID_ARGUMENT;
    e->lu_bits = bits;
    return BG_OK;
}

static int upload_field(DevBuf &b, const Field &f, const double **ptr, double *scalar) {
    *scalar = f.scalar;
    *ptr = nullptr;
    if (f.v.empty()) return BG_OK;
    int rc = b.ensure(f.v.size() * sizeof(double));
    if (rc != BG_OK) return rc;
    BG_HIP_CHECK(hipMemcpy(b.p, f.v.data(), f.v.size() * sizeof(double), hipMemcpyHostToDevice));
    *ptr = (const double *)b.p;
    return BG_OK;
}

static int ensemble_prepare(bg_ensemble *e, bg_ensemble_desc &d, bool want_order) {
    if (!e->dt_max.set || !e->dt_min.set || !e->tol.set || !e->t0.set || !e->t_end.set ||
        !e->have_y0 || !e->have_deriv)
        return BG_MISSING_PARAMETERS;
    memset(&d, 0, sizeof d);
    d.n = e->n;
    d.dim = e->dim;
    d.n_params = e->n_params;
    d.max_attempts = e->max_attempts;
    int rc;
#define UP(buf, vec, ptrfield)                                                               \
    if (!vec.empty()) {                                                                      \
        if ((rc = buf.ensure(vec.size() * sizeof(vec[0]))) != BG_OK) return rc;              \
        BG_HIP_CHECK(hipMemcpy(buf.p, vec.data(), vec.size() * sizeof(vec[0]),               \
                               hipMemcpyHostToDevice));                                      \
        d.ptrfield = (decltype(d.ptrfield))buf.p;                                            \
    }
    UP(e->d_y0, e->y0, y0)
    UP(e->d_params, e->params, params)
    UP(e->d_shared, e->shared, shared)
#undef UP
    if ((rc = upload_field(e->d_t0, e->t0, &d.t0, &d.t0_s)) != BG_OK) return rc;
    if ((rc = upload_field(e->d_tend, e->t_end, &d.t_end, &d.t_end_s)) != BG_OK) return rc;
    if ((rc = upload_field(e->d_tol, e->tol, &d.tol, &d.tol_s)) != BG_OK) return rc;
    if ((rc = upload_field(e->d_dtmin, e->dt_min, &d.dt_min, &d.dt_min_s)) != BG_OK) return rc;
    if ((rc = upload_field(e->d_dtmax, e->dt_max, &d.dt_max, &d.dt_max_s)) != BG_OK) return rc;
    if (want_order && e->n > 1 && (!e->t_end.v.empty() || !e->tol.v.empty() || !e->t0.v.empty())) {
        // longest-first schedule: predicted cost (t_end - t0) * tol^(-1/4)
        std::vector<double> cost(e->n);
        for (uint64_t i = 0; i < e->n; ++i)
            cost[i] = (e->t_end.at(i) - e->t0.at(i)) * pow(e->tol.at(i), -0.25);
        std::vector<int32_t> order(e->n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
        if ((rc = e->d_order.ensure(order.size() * sizeof(int32_t))) != BG_OK) return rc;
        BG_HIP_CHECK(hipMemcpy(e->d_order.p, order.data(), order.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice));
        d.order = (const int32_t *)e->d_order.p;
    }
    if ((rc = e->d_tout.ensure(e->n * sizeof(double))) != BG_OK) return rc;
    if ((rc = e->d_yout.ensure(e->n * e->dim * sizeof(double))) != BG_OK) return rc;
    if ((rc = e->d_status.ensure(e->n * sizeof(int32_t))) != BG_OK) return rc;
    if ((rc = e->d_counts.ensure(e->n * sizeof(bg_counts))) != BG_OK) return rc;
    if ((rc = e->d_work.ensure(e->n * sizeof(bg_counts))) != BG_OK) return rc;
    d.work = (bg_counts *)e->d_work.p;
    d.t_out = (double *)e->d_tout.p;
    d.y_out = (double *)e->d_yout.p;
    d.status = (int32_t *)e->d_status.p;
    d.counts = (bg_counts *)e->d_counts.p;
    return BG_OK;
}

extern "C" int bg_ensemble_solve(bg_ensemble *e) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    bg_ensemble_desc d;
    int rc = ensemble_prepare(e, d, true);
    if (rc != BG_OK) return rc;
    e->solved = false;
    if (is_rk(e->method)) {
        bg_launch lc;
        memset(&lc, 0, sizeof lc);
        rc = rk_launch_ensemble(e->method, e->rhs, e->semantics, d, lc, nullptr);
    } else if (is_adams(e->method)) {
        rc = adams_launch_ensemble(e->method, e->rhs, e->semantics, d, nullptr);
    } else {
        rc = bdf_launch_ensemble(e->method, e->rhs, e->semantics, e->lu_bits == 32, d, nullptr);
    }
    if (rc != BG_OK) return rc;
    BG_HIP_CHECK(hipStreamSynchronize(nullptr));
    e->solved = true;
    return BG_OK;
}

extern "C" int bg_ensemble_solve_plain(bg_ensemble *e) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    bg_ensemble_desc d;
    int rc = ensemble_prepare(e, d, false);
    if (rc != BG_OK) return rc;
    e->solved = false;
    if (!is_rk(e->method)) return BG_ERR_UNSUPPORTED;
    rc = rk_launch_plain(e->method, e->rhs, e->semantics, d, nullptr);
    if (rc != BG_OK) return rc;
    e->solved = true;
    return BG_OK;
}

// README.md:45-47: solve_ivp tries a fifth-order predictor-corrector, then
// Runge-Kutta-Fehlberg, then BDF6.  Batched: every instance starts with
// Adams5; the instances that did not finish OK are gathered into a new
// ensemble for RK45, and those still failing into one for BDF6.  A method
// that does not support the right-hand side is skipped.
static bool method_supports(int method, int rhs, uint32_t dim) {
    if (is_adams(method)) return adams_supported(method, rhs, dim);
    if (is_rk(method)) return rhs == BG_RHS_LINEAR ? linear_supported(dim) && method == BG_RK45
                                                   : rk_supported(method, rhs, dim);
    return bdf_supported(method, rhs, dim);
}

static bg_ensemble *ensemble_subset(const bg_ensemble *e, const std::vector<uint64_t> &idx,
                                    int method) {
    bg_ensemble *s = new bg_ensemble();
    const uint64_t m = idx.size(), n = e->n;
    s->method = method;
    s->dim = e->dim;
    s->n = m;
    s->semantics = e->semantics;
    s->rhs = e->rhs;
    s->n_params = e->n_params;
    s->max_attempts = e->max_attempts;
    s->lu_bits = e->lu_bits;
    auto sub = [&](const Field &f, Field &g) {
        g.set = f.set;
        g.scalar = f.scalar;
        if (!f.v.empty()) {
            g.v.resize(m);
            for (uint64_t j = 0; j < m; ++j) g.v[j] = f.v[idx[j]];
        }
    };
    sub(e->tol, s->tol);
    sub(e->dt_max, s->dt_max);
    sub(e->dt_min, s->dt_min);
    sub(e->t0, s->t0);
    sub(e->t_end, s->t_end);
    s->y0.resize((size_t)e->dim * m);
    for (uint32_t c = 0; c < e->dim; ++c)
        for (uint64_t j = 0; j < m; ++j) s->y0[c * m + j] = e->y0[c * n + idx[j]];
    s->params.resize((size_t)e->n_params * m);
    for (uint32_t k = 0; k < e->n_params; ++k)
        for (uint64_t j = 0; j < m; ++j) s->params[k * m + j] = e->params[k * n + idx[j]];
    s->shared = e->shared;
    s->have_y0 = e->have_y0;
    s->have_deriv = e->have_deriv;
    return s;
}

extern "C" int bg_ensemble_solve_ivp(bg_ensemble *e, int32_t *method_used) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    if (!e->have_y0 || !e->have_deriv || !e->tol.set || !e->dt_max.set || !e->dt_min.set ||
        !e->t0.set || !e->t_end.set)
        return BG_MISSING_PARAMETERS;
    const uint64_t n = e->n;
    const uint32_t D = e->dim;
    std::vector<double> t(n, 0.0), y((size_t)D * n);
    std::vector<int32_t> st(n, BG_ERR_UNSUPPORTED), used(n, -1);
    std::vector<bg_counts> cnt(n);
    memset(cnt.data(), 0, n * sizeof(bg_counts));
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t c = 0; c < D; ++c) y[c * n + i] = e->y0[c * n + i];
    std::vector<uint64_t> todo(n);
    for (uint64_t i = 0; i < n; ++i) todo[i] = i;
    const int chain[3] = {BG_ADAMS5, BG_RK45, BG_BDF6};
    for (int method : chain) {
        if (todo.empty()) break;
        if (!method_supports(method, e->rhs, D)) continue;
        bg_ensemble *s = ensemble_subset(e, todo, method);
        const uint64_t m = todo.size();
        std::vector<double> ts(m), ys((size_t)D * m);
        std::vector<int32_t> ss(m);
        std::vector<bg_counts> cs(m);
        int rc = bg_ensemble_solve(s);
        if (rc == BG_OK) rc = bg_ensemble_results(s, ts.data(), ys.data(), ss.data(), cs.data());
        delete s;
        if (rc != BG_OK) return rc;
        std::vector<uint64_t> next;
        for (uint64_t j = 0; j < m; ++j) {
            const uint64_t i = todo[j];
            t[i] = ts[j];
            for (uint32_t c = 0; c < D; ++c) y[c * n + i] = ys[c * m + j];
            st[i] = ss[j];
            cnt[i] = cs[j];
            used[i] = method;
            if (ss[j] != BG_OK) next.push_back(i);
        }
        todo.swap(next);
    }
    // results into the ensemble's own buffers
    int rc;
    if ((rc = e->d_tout.ensure(n * sizeof(double))) != BG_OK) return rc;
    if ((rc = e->d_yout.ensure((size_t)D * n * sizeof(double))) != BG_OK) return rc;
    if ((rc = e->d_status.ensure(n * sizeof(int32_t))) != BG_OK) return rc;
    if ((rc = e->d_counts.ensure(n * sizeof(bg_counts))) != BG_OK) return rc;
    if (n) {
        BG_HIP_CHECK(hipMemcpy(e->d_tout.p, t.data(), n * sizeof(double), hipMemcpyHostToDevice));
        BG_HIP_CHECK(hipMemcpy(e->d_yout.p, y.data(), (size_t)D * n * sizeof(double),
                               hipMemcpyHostToDevice));
        BG_HIP_CHECK(hipMemcpy(e->d_status.p, st.data(), n * sizeof(int32_t),
                               hipMemcpyHostToDevice));
        BG_HIP_CHECK(hipMemcpy(e->d_counts.p, cnt.data(), n * sizeof(bg_counts),
                               hipMemcpyHostToDevice));
    }
    if (method_used) memcpy(method_used, used.data(), n * sizeof(int32_t));
    e->solved = true;
    return BG_OK;
}

extern "C" int bg_ensemble_results(bg_ensemble *e, double *t_out, double *y_out,
                                   int32_t *status, bg_counts *counts) {
    if (!e) return BG_ERR_INVALID_ARGUMENT;
    if (!e->solved) return BG_ERR_STATE;
    if (e->n == 0) return BG_OK;
    if (t_out)
        BG_HIP_CHECK(hipMemcpy(t_out, e->d_tout.p, e->n * sizeof(double), hipMemcpyDeviceToHost));
    if (y_out)
        BG_HIP_CHECK(hipMemcpy(y_out, e->d_yout.p, e->n * e->dim * sizeof(double),
                               hipMemcpyDeviceToHost));
    if (status)
        BG_HIP_CHECK(
            hipMemcpy(status, e->d_status.p, e->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (counts)
        BG_HIP_CHECK(
            hipMemcpy(counts, e->d_counts.p, e->n * sizeof(bg_counts), hipMemcpyDeviceToHost));
    return BG_OK;
}

extern "C" int bg_ensemble_work(bg_ensemble *e, bg_counts *work) {
    if (!e || !work) return BG_ERR_INVALID_ARGUMENT;
    if (!e->solved) return BG_ERR_STATE;
    if (e->n == 0) return BG_OK;
    BG_HIP_CHECK(hipMemcpy(work, e->d_work.p, e->n * sizeof(bg_counts), hipMemcpyDeviceToHost));
    return BG_OK;
}

extern "C" void bg_ensemble_destroy(bg_ensemble *e) { delete e; }

// ===========================================================================
// device-pointer entry points
// ===========================================================================
extern "C" int bg_rk_solve_device(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                                  const bg_launch *launch, void *stream) {
    if (!d || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    bg_launch lc;
    memset(&lc, 0, sizeof lc);
    if (launch) lc = *launch;
    return rk_launch_ensemble(method, rhs, semantics, *d, lc, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Allocation-free RK launch.  The persistent launch's workspace (the work
// counter at [0, 256), the descriptor copy after it) lives in this file-scope
// device array instead of coming from the stream-ordered allocator.  Each
// GPU holds its own copy until the process ends; its address is looked up
// once per device (hipGetSymbolAddress, with that device current) and only
// that address is handed to kernels.  The setup kernel rewrites the counter
// and the descriptor at the start of every launch, so nothing carries over
// from the previous call; calls that overlap on one GPU share the array and
// race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 1024;
static_assert(kLaunchWorkspaceBytes <= kNoAllocScratchBytes,
              "the RK launch workspace must fit the allocation-free scratch");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static void *g_scratch_addr[64];

// scratch's device address on the current device (nullptr on failure)
static void *noalloc_scratch(int dev) {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_addr[dev] = p;
    }
    return g_scratch_addr[dev];
}

extern "C" int bg_rk_solve_device_noalloc(int method, int rhs, int semantics,
                                          const bg_ensemble_desc *d, const bg_launch *launch,
                                          void *stream) {
    if (!d || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    int dev = 0;
    BG_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return BG_ERR_UNSUPPORTED;
    void *ws = noalloc_scratch(dev);
    if (!ws) return BG_ERR_HIP;
    bg_launch lc;
    memset(&lc, 0, sizeof lc);
    if (launch) lc = *launch;
    return rk_launch_ensemble(method, rhs, semantics, *d, lc, (hipStream_t)stream, ws);
}

extern "C" int bg_rk_solve_path_device(int method, int rhs, int semantics,
                                       const bg_ensemble_desc *d, const bg_path_desc *path,
                                       void *stream) {
    if (!d || !path || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    return rk_launch_path_ensemble(method, rhs, semantics, *d, *path, (hipStream_t)stream);
}

extern "C" int bg_rk_autotune_device(int method, int rhs, int semantics,
                                     const bg_ensemble_desc *d, bg_tuning *out, void *stream) {
    if (!d || !out || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    return rk_autotune(method, rhs, semantics, *d, *out, (hipStream_t)stream);
}

extern "C" int bg_rk_schedule_device(const bg_tuning *tuning, const bg_ensemble_desc *d,
                                     int32_t *order, void *stream) {
    if (!d || !tuning) return BG_ERR_INVALID_ARGUMENT;
    return rk_schedule(*tuning, *d, order, (hipStream_t)stream);
}

extern "C" int bg_schedule_by_counts_device(const bg_counts *counts, uint64_t n, int32_t *order,
                                            void *stream) {
    return schedule_by_counts(counts, n, order, (hipStream_t)stream);
}

extern "C" int bg_rk_solve_tuned_device(int method, int rhs, int semantics,
                                        const bg_ensemble_desc *d, const bg_tuning *tuning,
                                        void *stream) {
    if (!d || !tuning || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    return rk_solve_tuned(method, rhs, semantics, *d, *tuning, (hipStream_t)stream);
}

extern "C" int bg_rk_solve_plain_device(int method, int rhs, int semantics,
                                        const bg_ensemble_desc *d, void *stream) {
    if (!d || !is_rk(method)) return BG_ERR_INVALID_ARGUMENT;
    return rk_launch_plain(method, rhs, semantics, *d, (hipStream_t)stream);
}

extern "C" int bg_bdf_solve_device(int method, int rhs, int semantics, int lu_bits,
                                   const bg_ensemble_desc *d, void *stream) {
    if (!d || !is_bdf(method) || (lu_bits != 32 && lu_bits != 64)) return BG_ERR_INVALID_ARGUMENT;
    return bdf_launch_ensemble(method, rhs, semantics, lu_bits == 32, *d, (hipStream_t)stream);
}

extern "C" int bg_bdf_autotune_device(int method, int rhs, int semantics, int lu_bits,
                                      const bg_ensemble_desc *d, bg_tuning *out, void *stream) {
    if (!d || !out || !is_bdf(method) || (lu_bits != 32 && lu_bits != 64))
        return BG_ERR_INVALID_ARGUMENT;
    return bdf_autotune(method, rhs, semantics, lu_bits == 32, *d, *out, (hipStream_t)stream);
}

extern "C" int bg_bdf_solve_tuned_device(int method, int rhs, int semantics, int lu_bits,
                                         const bg_ensemble_desc *d, const bg_tuning *tuning,
                                         void *stream) {
    if (!d || !tuning || !is_bdf(method) || (lu_bits != 32 && lu_bits != 64))
        return BG_ERR_INVALID_ARGUMENT;
    return bdf_solve_tuned(method, rhs, semantics, lu_bits == 32, *d, *tuning,
                           (hipStream_t)stream);
}

extern "C" int bg_bdf_inverse_device(uint32_t dim, uint64_t count, const double *a, double *inv,
                                     int32_t *branch, int first, void *stream) {
    return bdf_launch_inverse(dim, count, a, inv, branch, first, (hipStream_t)stream);
}

extern "C" int bg_adams_solve_device(int method, int rhs, int semantics,
                                     const bg_ensemble_desc *d, void *stream) {
    if (!d || !is_adams(method)) return BG_ERR_INVALID_ARGUMENT;
    return adams_launch_ensemble(method, rhs, semantics, *d, (hipStream_t)stream);
}

extern "C" int bg_rk_vjp_device(int method, int rhs, const bg_adjoint_desc *d, void *stream) {
    if (!d || method != BG_RK45) return BG_ERR_INVALID_ARGUMENT;
    return rk_launch_adjoint(method, rhs, *d, (hipStream_t)stream);
}

extern "C" int bg_adjoint_solve_tuned_device(int method, int rhs, const bg_adjoint_desc *d,
                                             const bg_adjoint_tuning *plan, void *stream) {
    if (!d || !plan || method != BG_RK45) return BG_ERR_INVALID_ARGUMENT;
    return adjoint_solve_tuned(method, rhs, *d, *plan, (hipStream_t)stream);
}

extern "C" int bg_adjoint_autotune_device(int method, int rhs, const bg_adjoint_desc *d,
                                          bg_adjoint_tuning *out, void *stream) {
    if (!d || !out || method != BG_RK45) return BG_ERR_INVALID_ARGUMENT;
    return adjoint_autotune(method, rhs, *d, *out, (hipStream_t)stream);
}

extern "C" int bg_version(void) { return 1; }

extern "C" int bg_device_info(int *n_cu, int *arch_major, int *arch_minor) {
    int dev = 0;
    BG_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    BG_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (arch_major) *arch_major = prop.major;
    if (arch_minor) *arch_minor = prop.minor;
    return BG_OK;
}
