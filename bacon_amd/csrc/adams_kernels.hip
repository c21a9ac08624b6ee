// adams_kernels.hip — batched Adams-Bashforth-Moulton predictor-corrector
// (Adams5 / Adams3) for gfx950.
//
// Reference: bacon-sci's AdamsSolver (src/ivp/adams.rs:336-633): an
// (O-1)-step Adams-Bashforth predictor, an O-coefficient Adams-Moulton
// corrector evaluated once (PECE), the error estimate
// (19/270) / dt * |corrector - predictor|, an RK4 bootstrap of O - 1 steps
// after every step-size change, and the yield bookkeeping of the iterator.
//
// One lane per instance, refilled from a global work counter like the RK
// ensemble kernel.  The derivative deque holds at most O - 1 entries when the
// predictor runs, so it lives in registers as a shift register (a push that
// follows a full deque drops the oldest entry, which is the reference's
// push_back + pop_front).  The value deque only matters through its length
// here (final states, not the path).  Arithmetic follows the oracle's DEVICE
// profile operation by operation (explicit fma, rootn() for the step-size
// factor), so results are bit-identical to it.
//
// Semantics: as built, the derivative of the first Adams step after a
// bootstrap never joins the deque (adams.rs:434-441 pushes the value only),
// leaving the history one step stale; intended pushes it.
#include "bg_device.h"
#include "bg_internal.h"

namespace bg {

static constexpr uint32_t kAdamsDefaultMaxSteps = 10000000u;

// x^(1/N), N = 3 or 5, from exact operations (frexp, fma, ldexp, rounded
// constants): x = m 2^e with m in [1, 2); m^(1/N) by a degree-6 polynomial in
// (m - 1.5) interpolating at Chebyshev nodes (relative error < 2.5e-7);
// 2^(e/N) = 2^floor(e/N) * 2^((e mod N)/N).  0, inf and NaN map to themselves.
template <int N>
__device__ __forceinline__ double rootn(double x) {
    static_assert(N == 3 || N == 5, "rootn: N = 3 or 5");
    if (!(x == x) || x == 0.0 || x == __builtin_inf()) return x;
    int ex;
    const double m = frexp(x, &ex);
    const double u = fma(m, 2.0, -1.5);
    const int e = ex - 1;
    double p;
    if constexpr (N == 3) {
        p = kc(-0.0027033338703399646);
        p = fma(p, u, kc(0.0051666994182390956));
        p = fma(p, u, kc(-0.009260767947395939));
        p = fma(p, u, kc(0.02085714929540036));
        p = fma(p, u, kc(-0.05653050145457822));
        p = fma(p, u, kc(0.254383453864114));
        p = fma(p, u, kc(1.1447142425533323));
    } else {
        p = kc(-0.002242595967870246);
        p = fma(p, u, kc(0.004170036250587015));
        p = fma(p, u, kc(-0.007159250830014004));
        p = fma(p, u, kc(0.015356114792636277));
        p = fma(p, u, kc(-0.038560210159460984));
        p = fma(p, u, kc(0.14459836703782772));
        p = fma(p, u, kc(1.0844717711976986));
    }
    int a = e / N, b = e % N;
    if (b < 0) {
        b += N;
        a -= 1;
    }
    double s;
    if constexpr (N == 3)
        s = b == 0 ? 1.0 : (b == 1 ? 1.2599210498948732 : 1.5874010519681994);
    else
        s = b == 0 ? 1.0
                   : (b == 1 ? 1.148698354997035
                             : (b == 2 ? 1.3195079107728942
                                       : (b == 3 ? 1.515716566510398 : 1.7411011265922482)));
    return ldexp(p * s, a);
}

// adams.rs:568-661
template <int O>
struct AdamsCoef;
template <>
struct AdamsCoef<5> {
    __device__ static constexpr double pred(int k) {
        constexpr double tf = 24.0;
        constexpr double c[5] = {55.0 / tf, -59.0 / tf, 37.0 / tf, -9.0 / tf, 0.0};
        return c[k];
    }
    __device__ static constexpr double corr(int k) {
        constexpr double sh = 720.0;
        constexpr double c[5] = {251.0 / sh, 646.0 / sh, -264.0 / sh, 106.0 / sh, -19.0 / sh};
        return c[k];
    }
};
template <>
struct AdamsCoef<3> {
    __device__ static constexpr double pred(int k) {
        constexpr double c[3] = {1.0 + 1.0 / 2.0, -(1.0 / 2.0), 0.0};
        return c[k];
    }
    __device__ static constexpr double corr(int k) {
        constexpr double c[3] = {5.0 / 12.0, 2.0 / 3.0, -(1.0 / 12.0)};
        return c[k];
    }
};

enum : int { ADAMS_YIELD = -1, ADAMS_REDO = -2, ADAMS_DONE = -3 };

// PATH: also keep the values of the last bootstrap (the value deque's
// entries, oldest first), which the yield_memory countdown of adams.rs:415-428
// yields; `yitem` names what a yield returns: -1 the current (t, y), else
// that entry.  Without PATH both are dead and compile away.
template <class RHS, int O, bool AsBuilt, bool PATH = false>
struct AdamsLane {
    static constexpr int D = RHS::D;
    using C = AdamsCoef<O>;
    double t, dt, end, tol, dt_min, dt_max;
    double y[D], save[D], impl[D];
    double dq[O - 1][D];  // derivative history, oldest first (shift register)
    double vt[PATH ? O - 1 : 1], vy[PATH ? O - 1 : 1][D];  // value deque (PATH only)
    int vlen;             // length of the value deque
    int ym;               // yield_memory
    int yitem;
    double p[RHS::P > 0 ? RHS::P : 1];
    uint32_t n_rej;

    __device__ __forceinline__ bool f(double tt, const double (&yy)[D], double (&out)[D]) const {
        return RHS::eval(tt, yy, p, out);
    }
    __device__ __forceinline__ void push_d(const double (&v)[D]) {
#pragma unroll
        for (int k = 0; k < O - 2; ++k)
#pragma unroll
            for (int c = 0; c < D; ++c) dq[k][c] = dq[k + 1][c];
#pragma unroll
        for (int c = 0; c < D; ++c) dq[O - 2][c] = v[c];
    }
    // value push at the full deque: push_back + pop_front, as for dq
    __device__ __forceinline__ void push_v() {
        if constexpr (PATH) {
#pragma unroll
            for (int k = 0; k < O - 2; ++k) {
                vt[k] = vt[k + 1];
#pragma unroll
                for (int c = 0; c < D; ++c) vy[k][c] = vy[k + 1][c];
            }
            vt[O - 2] = t;
#pragma unroll
            for (int c = 0; c < D; ++c) vy[O - 2][c] = y[c];
        }
    }

    // adams.rs:344-396
    template <int IT>
    __device__ __forceinline__ int runge_kutta() {
#pragma unroll 1
        for (int it = 0; it < IT; ++it) {
            double k1[D], k2[D], k3[D], k4[D], mid[D];
            if (!f(t, y, k1)) return BG_USER_ERROR;
#pragma unroll
            for (int c = 0; c < D; ++c) k1[c] *= dt;
#pragma unroll
            for (int c = 0; c < D; ++c) mid[c] = fma(k1[c], 0.5, y[c]);
            if (!f(fma(0.5, dt, t), mid, k2)) return BG_USER_ERROR;
#pragma unroll
            for (int c = 0; c < D; ++c) k2[c] *= dt;
#pragma unroll
            for (int c = 0; c < D; ++c) mid[c] = fma(k2[c], 0.5, y[c]);
            if (!f(fma(0.5, dt, t), mid, k3)) return BG_USER_ERROR;
#pragma unroll
            for (int c = 0; c < D; ++c) k3[c] *= dt;
#pragma unroll
            for (int c = 0; c < D; ++c) mid[c] = y[c] + k3[c];
            if (!f(t + dt, mid, k4)) return BG_USER_ERROR;
#pragma unroll
            for (int c = 0; c < D; ++c) k4[c] *= dt;
            if (it != 0) {
                double fd[D];
                if (!f(t, y, fd)) return BG_USER_ERROR;
                push_d(fd);
                push_v();
                ++vlen;
            }
#pragma unroll
            for (int c = 0; c < D; ++c) {
                double sum = fma(k2[c], 2.0, k1[c]);
                sum = fma(k3[c], 2.0, sum);
                sum = sum + k4[c];
                y[c] = fma(sum, kc(1.0 / 6.0), y[c]);
            }
            t += dt;
        }
        double fd[D];
        if (!f(t, y, fd)) return BG_USER_ERROR;
        push_d(fd);
        push_v();
        ++vlen;
        return BG_OK;
    }

    __device__ __forceinline__ double q_of(double error) const {
        return rootn<O>(tol / (2.0 * error));
    }

    // One AdamsSolver::step (adams.rs:411-561)
    __device__ __forceinline__ int step() {
        yitem = -1;
        if (ym > 0 && ym < O) {  // adams.rs:415-428
            yitem = O - ym - 1;
            ym -= 1;
            if (ym == 0) ym = O + 1;
            return ADAMS_YIELD;
        }
        if (ym == O + 1) {  // adams.rs:434-441
            ym = 0;
            if (!AsBuilt) push_d(impl);
            return ADAMS_YIELD;
        }
        if (t >= end) return ADAMS_DONE;
        if (t + dt >= end) {  // adams.rs:447-451
            dt = end - t;
            const int rc = runge_kutta<1>();
            if (rc) return rc;
            return ADAMS_YIELD;
        }
        if (vlen == 0) {  // adams.rs:453-464
#pragma unroll
            for (int c = 0; c < D; ++c) save[c] = y[c];
            if (t + dt * (double)(O - 1) >= end) dt = (end - t) / (double)(O - 1);
            const int rc = runge_kutta<O - 1>();
            if (rc) return rc;
            ym = O;
            return ADAMS_REDO;
        }
        // adams.rs:466-486
        double pr[D], co[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double sc = dq[0][c] * kc(C::pred(O - 2));
#pragma unroll
            for (int i = 1; i < O - 1; ++i) sc = fma(dq[i][c], kc(C::pred(O - i - 2)), sc);
            pr[c] = fma(sc, dt, y[c]);
        }
        if (!f(t + dt, pr, impl)) return BG_USER_ERROR;
        double ss = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double sc = impl[c] * kc(C::corr(0));
#pragma unroll
            for (int i = 0; i < O - 1; ++i) sc = fma(dq[i][c], kc(C::corr(O - i - 1)), sc);
            co[c] = fma(sc, dt, y[c]);
            const double df = co[c] - pr[c];
            ss = c == 0 ? df * df : fma(df, df, ss);
        }
        const double error = (kc(19.0 / 270.0) / dt) * sqrt(ss);
        if (error <= tol) {  // adams.rs:490-525
#pragma unroll
            for (int c = 0; c < D; ++c) y[c] = co[c];
            t += dt;
            if (ym == O) {
                ym -= 1;
                return ADAMS_REDO;
            }
            push_d(impl);  // with the value push: lengths stay O - 1
            if (error < 0.1 * tol) {
                const double q = q_of(error);
                if (q > 4.0)
                    dt *= 4.0;
                else
                    dt *= q;
                if (dt > dt_max) dt = dt_max;
                vlen = 0;
            }
            return ADAMS_YIELD;
        }
        ++n_rej;
        if (ym == O) {  // adams.rs:531-536
            t -= dt * (double)(O - 1);
#pragma unroll
            for (int c = 0; c < D; ++c) y[c] = save[c];
        }
        const double q = q_of(error);  // adams.rs:538-556
        if (q < 0.1)
            dt *= 0.1;
        else
            dt *= q;
        if (dt < dt_min) return BG_MINIMUM_TIME_DELTA_EXCEEDED;
        vlen = 0;
        return ADAMS_REDO;
    }

    __device__ __forceinline__ void load(const bg_ensemble_desc &d, uint64_t i) {
        const uint64_t n = d.n;
#pragma unroll
        for (int c = 0; c < D; ++c) y[c] = d.y0[(uint64_t)c * n + i];
#pragma unroll
        for (int k = 0; k < RHS::P; ++k) p[k] = d.params[(uint64_t)k * n + i];
        t = d.t0 ? d.t0[i] : d.t0_s;
        end = d.t_end ? d.t_end[i] : d.t_end_s;
        tol = d.tol ? d.tol[i] : d.tol_s;
        dt_min = d.dt_min ? d.dt_min[i] : d.dt_min_s;
        dt_max = d.dt_max ? d.dt_max[i] : d.dt_max_s;
        dt = (dt_max + dt_min) * 0.5;  // adams.rs:288
        vlen = 0;
        ym = 0;
        n_rej = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) save[c] = impl[c] = 0.0;
#pragma unroll
        for (int k = 0; k < O - 1; ++k)
#pragma unroll
            for (int c = 0; c < D; ++c) dq[k][c] = 0.0;
    }
};

template <class RHS, int O, bool AsBuilt>
__global__ void __launch_bounds__(256)
    adams_ensemble_kernel(const bg_ensemble_desc *dv, uint32_t *__restrict__ counter) {
    constexpr int D = RHS::D;
    AdamsLane<RHS, O, AsBuilt> L;
    const int lane = __lane_id();
    const uint64_t max_att = dv->max_attempts;
    const uint32_t cap =
        (uint32_t)((max_att && max_att < kAdamsDefaultMaxSteps) ? max_att : kAdamsDefaultMaxSteps);
    int64_t inst = -1;  // -1 need work, -2 exhausted
    uint32_t steps = 0, nacc = 0;
    for (;;) {
        const uint64_t need = __ballot(inst == -1);
        if (need != 0) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (inst == -1) {
                const uint64_t pos =
                    (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
                const bg_ensemble_desc di = desc_inputs(dv);
                if (pos < di.n) {
                    inst = di.order ? (int64_t)di.order[pos] : (int64_t)pos;
                    L.load(di, (uint64_t)inst);
                    steps = nacc = 0;
                } else {
                    inst = -2;
                }
            }
        }
        if (__ballot(inst >= 0) == 0) break;
        int st = -1;
        for (;;) {
            if (__ballot(inst >= 0 && st >= 0) != 0) break;
            if (inst >= 0 && st < 0) {
                if (steps >= cap) {
                    st = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                } else {
                    ++steps;
                    const int r = L.step();
                    if (r == ADAMS_YIELD)
                        ++nacc;
                    else if (r == ADAMS_DONE)
                        st = BG_OK;
                    else if (r != ADAMS_REDO)
                        st = r;
                }
            }
        }
        if (inst >= 0 && st >= 0) {
            const bg_ensemble_desc d = desc_outputs(dv);
            const uint64_t n = d.n, i = (uint64_t)inst;
            if (d.t_out) d.t_out[i] = L.t;
            if (d.y_out) {
#pragma unroll
                for (int c = 0; c < D; ++c) d.y_out[(uint64_t)c * n + i] = L.y[c];
            }
            if (d.status) d.status[i] = st;
            if (d.counts) {
                bg_counts cn;
                cn.n_accepted = nacc;
                cn.n_rejected = L.n_rej;
                d.counts[i] = cn;
            }
            inst = -1;
        }
    }
}

// Single-instance path production: every yield of the iterator (adams.rs:
// 398-565 through ivp.rs:205-239) into the path buffer; stops when the
// buffer is full (resumable: the lane state goes back to *st) or the solve
// ends.  Steps are counted like the ensemble kernel's, as step() calls.
template <class RHS, int O, bool AsBuilt>
__global__ void __launch_bounds__(64)
    adams_path_kernel(AdamsPathState *__restrict__ st, double *__restrict__ path_t,
                      double *__restrict__ path_y, uint32_t cap) {
    constexpr int D = RHS::D;
    static_assert(D <= 4 && O - 1 <= 4 && RHS::P <= 8, "AdamsPathState sizes");
    if (threadIdx.x != 0) return;
    AdamsLane<RHS, O, AsBuilt, true> L;
    L.t = st->t;
    L.dt = st->dt;
    L.end = st->end;
    L.tol = st->tol;
    L.dt_min = st->dt_min;
    L.dt_max = st->dt_max;
    L.vlen = st->vlen;
    L.ym = st->ym;
    L.yitem = -1;
    L.n_rej = st->n_rejected;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        L.y[c] = st->y[c];
        L.save[c] = st->save[c];
        L.impl[c] = st->impl[c];
    }
#pragma unroll
    for (int k = 0; k < O - 1; ++k) {
        L.vt[k] = st->vt[k];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            L.dq[k][c] = st->dq[k][c];
            L.vy[k][c] = st->vy[k][c];
        }
    }
#pragma unroll
    for (int k = 0; k < RHS::P; ++k) L.p[k] = st->p[k];
    uint64_t steps = st->steps;
    const uint64_t max_steps = st->max_steps;
    uint32_t nacc = st->n_accepted, count = 0;
    int status = st->status, finished = st->finished;
    while (!finished && count < cap) {
        if (steps >= max_steps) {
            status = BG_MAXIMUM_ITERATIONS_EXCEEDED;
            finished = 1;
            break;
        }
        ++steps;
        const int r = L.step();
        if (r == ADAMS_YIELD) {
            ++nacc;
            double yt = L.t, yv[D];
#pragma unroll
            for (int c = 0; c < D; ++c) yv[c] = L.y[c];
#pragma unroll
            for (int k = 0; k < O - 1; ++k) {
                if (L.yitem == k) {
                    yt = L.vt[k];
#pragma unroll
                    for (int c = 0; c < D; ++c) yv[c] = L.vy[k][c];
                }
            }
            path_t[count] = yt;
#pragma unroll
            for (int c = 0; c < D; ++c) path_y[(uint64_t)count * D + c] = yv[c];
            ++count;
        } else if (r == ADAMS_DONE) {
            finished = 1;
        } else if (r != ADAMS_REDO) {
            status = r;
            finished = 1;
        }
    }
    st->t = L.t;
    st->dt = L.dt;
    st->vlen = L.vlen;
    st->ym = L.ym;
    st->n_rejected = L.n_rej;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        st->y[c] = L.y[c];
        st->save[c] = L.save[c];
        st->impl[c] = L.impl[c];
    }
#pragma unroll
    for (int k = 0; k < O - 1; ++k) {
        st->vt[k] = L.vt[k];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            st->dq[k][c] = L.dq[k][c];
            st->vy[k][c] = L.vy[k][c];
        }
    }
    st->steps = steps;
    st->n_accepted = nacc;
    st->status = status;
    st->finished = finished;
    st->count = count;
}

__global__ void adams_setup_kernel(uint32_t *counter, bg_ensemble_desc *dst,
                                   const bg_ensemble_desc d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *dst = d;
    }
}

template <class T>
struct AdTag {
    using type = T;
};

template <class V>
static int visit_adams(int method, int rhs, int semantics, uint32_t dim, V &&v) {
    if (method != BG_ADAMS5 && method != BG_ADAMS3) return BG_ERR_INVALID_ARGUMENT;
    if (semantics != BG_AS_BUILT && semantics != BG_INTENDED) return BG_ERR_INVALID_ARGUMENT;
    auto with = [&](auto rtag) -> int {
        using R = typename decltype(rtag)::type;
        if (dim != (uint32_t)R::D) return BG_ERR_INVALID_ARGUMENT;
        const bool ab = semantics == BG_AS_BUILT;
        if (method == BG_ADAMS5)
            return ab ? v(rtag, std::integral_constant<int, 5>{}, std::true_type{})
                      : v(rtag, std::integral_constant<int, 5>{}, std::false_type{});
        return ab ? v(rtag, std::integral_constant<int, 3>{}, std::true_type{})
                  : v(rtag, std::integral_constant<int, 3>{}, std::false_type{});
    };
    switch (rhs) {
    case BG_RHS_EXP: return with(AdTag<RhsExp<1>>{});
    case BG_RHS_NEG: return with(AdTag<RhsNeg<1>>{});
    case BG_RHS_QUADRATIC: return with(AdTag<RhsQuadratic<1>>{});
    case BG_RHS_COS: return with(AdTag<RhsCos<1>>{});
    case BG_RHS_LORENZ: return with(AdTag<RhsLorenz>{});
    case BG_RHS_LOTKA_VOLTERRA: return with(AdTag<RhsLotkaVolterra>{});
    case BG_RHS_OSCILLATOR: return with(AdTag<RhsOscillator>{});
    case BG_RHS_ROBERTSON: return with(AdTag<RhsRobertson>{});
    case BG_RHS_TWO_LEVEL: return with(AdTag<RhsTwoLevel>{});
    default: return BG_ERR_UNSUPPORTED;
    }
}

bool adams_supported(int method, int rhs, uint32_t dim) {
    return visit_adams(method, rhs, BG_INTENDED, dim, [](auto, auto, auto) { return BG_OK; }) ==
           BG_OK;
}

int adams_launch_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                          hipStream_t stream, void *ws_in) {
    return visit_adams(method, rhs, semantics, d.dim, [&](auto rtag, auto otag, auto atag) -> int {
        using R = typename decltype(rtag)::type;
        constexpr int O = decltype(otag)::value;
        constexpr bool AB = decltype(atag)::value;
        if (d.n == 0) return BG_OK;
        if (d.n > 0x7fffffffull || !d.y0) return BG_ERR_INVALID_ARGUMENT;
        if (R::P > 0 && (!d.params || d.n_params < (uint32_t)R::P))
            return BG_ERR_INVALID_ARGUMENT;
        int dev = 0;
        BG_HIP_CHECK(hipGetDevice(&dev));
        int per_cu = 0;
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, adams_ensemble_kernel<R, O, AB>, 256, 0));
        // all resident blocks: 6 % faster than two per CU on 2^18 and 2^20
        // Lorenz-63 instances (register-limited to 3 waves per SIMD)
        if (per_cu < 1) per_cu = 1;
        uint64_t grid = (uint64_t)per_cu * (uint64_t)device_cu_count(dev);
        const uint64_t need = (d.n + 255) / 256;
        if (grid > need) grid = need;
        char *ws = (char *)ws_in;
        if (!ws_in) BG_HIP_CHECK(hipMallocAsync((void **)&ws, kLaunchWorkspaceBytes, stream));
        uint32_t *counter = (uint32_t *)ws;
        bg_ensemble_desc *ddesc = (bg_ensemble_desc *)(ws + 256);
        adams_setup_kernel<<<1, 64, 0, stream>>>(counter, ddesc, d);
        adams_ensemble_kernel<R, O, AB><<<(uint32_t)grid, 256, 0, stream>>>(ddesc, counter);
        const hipError_t launched = hipGetLastError();
        if (!ws_in) BG_HIP_CHECK(hipFreeAsync(ws, stream));
        BG_HIP_CHECK(launched);
        return BG_OK;
    });
}

int adams_launch_path(int method, int rhs, int semantics, uint32_t dim, AdamsPathState *state,
                      double *path_t, double *path_y, uint32_t cap, hipStream_t stream) {
    return visit_adams(method, rhs, semantics, dim, [&](auto rtag, auto otag, auto atag) -> int {
        using R = typename decltype(rtag)::type;
        constexpr int O = decltype(otag)::value;
        constexpr bool AB = decltype(atag)::value;
        adams_path_kernel<R, O, AB><<<1, 64, 0, stream>>>(state, path_t, path_y, cap);
        BG_HIP_CHECK(hipGetLastError());
        return BG_OK;
    });
}

}  // namespace bg
