// rk_kernels.hip — batched embedded Runge-Kutta (RK45 Fehlberg / RK23
// Bogacki-Shampine) for gfx950.
//
// Hot path: bacon-sci's RungeKuttaSolver::step (src/ivp/rk.rs:361-423) driven
// by IVPIterator::next (src/ivp.rs:220-238), recast as an ensemble of
// independent instances.
//
// Layout and schedule:
//  * one lane owns one instance; its state (t, dt, y[D], K[S][D], params)
//    lives in VGPRs for the whole solve;
//  * inputs and outputs are structure-of-arrays in HBM ([component][n]);
//  * the launch is persistent (grid = resident blocks); a wave refills its
//    idle lanes from a global work counter with one atomicAdd per wave
//    (ballot + popcount + mbcnt rank), so finished instances retire without
//    stalling the other lanes;
//  * an optional schedule permutation hands out instances longest-first.
#include "bg_device.h"
#include "bg_internal.h"

namespace bg {

static constexpr uint64_t kDefaultMaxAttempts = 10000000ull;
static constexpr int64_t kNeedWork = -1;
static constexpr int64_t kExhausted = -2;

template <class T>
struct Tag {
    using type = T;
};

template <class RHS, class Tab, bool H>
__device__ __forceinline__ void rk_load(RkLane<RHS, Tab, H> &L, const bg_ensemble_desc &d,
                                        uint64_t i) {
    constexpr int D = RHS::D, S = Tab::S;
    const uint64_t n = d.n;
#pragma unroll
    for (int k = 0; k < D; ++k) L.y[k] = d.y0[k * n + i];
#pragma unroll
    for (int k = 0; k < RHS::P; ++k) L.p[k] = d.params[k * n + i];
    L.t = d.t0 ? d.t0[i] : d.t0_s;
    L.end = d.t_end ? d.t_end[i] : d.t_end_s;
    L.tol = d.tol ? d.tol[i] : d.tol_s;
    L.dt_min = d.dt_min ? d.dt_min[i] : d.dt_min_s;
    L.dt_max = d.dt_max ? d.dt_max[i] : d.dt_max_s;
    L.dt = (L.dt_max + L.dt_min) * 0.5;  // rk.rs:315
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < D; ++k) L.K[s][k] = 0.0;  // rk.rs:323-327
}

template <class RHS, class Tab, bool H>
__device__ __forceinline__ void rk_store(const RkLane<RHS, Tab, H> &L, const bg_ensemble_desc &d,
                                         uint64_t i, int status, uint32_t nacc, uint32_t nrej) {
    constexpr int D = RHS::D;
    const uint64_t n = d.n;
    if (d.t_out) d.t_out[i] = L.t;
    if (d.y_out) {
#pragma unroll
        for (int k = 0; k < D; ++k) d.y_out[k * n + i] = L.y[k];
    }
    if (d.status) d.status[i] = status;
    if (d.counts) {
        bg_counts c;
        c.n_accepted = nacc;
        c.n_rejected = nrej;
        d.counts[i] = c;
    }
}

// One attempt plus the iterator's bookkeeping.  Returns -1 to continue, or
// the final status (BG_OK at Done).
template <class RHS, class Tab, bool H>
__device__ __forceinline__ int rk_advance(RkLane<RHS, Tab, H> &L, uint32_t &attempts, uint32_t cap,
                                          uint32_t &nacc, uint32_t &nrej) {
    if (attempts >= cap) return BG_MAXIMUM_ITERATIONS_EXCEEDED;
    const int st = L.attempt();
    ++attempts;
    if (st == STEP_ACCEPTED) {
        ++nacc;
        // the iterator's next step() call would return Done (rk.rs:362-364)
        return L.t >= L.end ? BG_OK : -1;
    }
    if (st == STEP_REJECTED) {
        ++nrej;
        return -1;
    }
    return st;
}

// Persistent ensemble kernel.  Outer loop: refill idle lanes (one atomicAdd
// per wave), store finished instances.  Inner loop: step attempts until at
// least `refill_min` lanes of the wave have finished (or all have).  Waves
// holding an instance among the first `prio_cutoff` schedule positions run at
// raised wave priority so the longest solves are not starved.
template <class RHS, class Tab>
__global__ void __launch_bounds__(256)
    rk_ensemble_kernel(const bg_ensemble_desc *dv, uint32_t *__restrict__ counter,
                       uint32_t refill_min, uint32_t prio_cutoff) {
    r8_init();
    RkLane<RHS, Tab, true> L;
    const int lane = __lane_id();
    const uint64_t max_att = dv->max_attempts;
    const uint32_t cap = (uint32_t)((max_att && max_att < kDefaultMaxAttempts) ? max_att
                                                                                : kDefaultMaxAttempts);
    int64_t inst = kNeedWork;
    uint64_t pos = 0;
    uint32_t attempts = 0, nacc = 0, nrej = 0;
    for (;;) {
        // ---- refill --------------------------------------------------------
        const uint64_t need = __ballot(inst == kNeedWork);
        if (need != 0) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (inst == kNeedWork) {
                const uint64_t below = need & ((1ull << lane) - 1ull);
                pos = (uint64_t)base + (uint64_t)__popcll(below);
                const bg_ensemble_desc di = desc_inputs(dv);  // one batch of loads
                if (pos < di.n) {
                    inst = di.order ? (int64_t)di.order[pos] : (int64_t)pos;
                    rk_load(L, di, (uint64_t)inst);
                    attempts = nacc = nrej = 0;
                } else {
                    inst = kExhausted;
                }
            }
            if (prio_cutoff) {
                if (__ballot(inst >= 0 && pos < prio_cutoff) != 0)
                    __builtin_amdgcn_s_setprio(2);
                else
                    __builtin_amdgcn_s_setprio(0);
            }
        }
        if (__ballot(inst >= 0) == 0) break;
        // ---- attempts ------------------------------------------------------
        const uint64_t active = __ballot(inst >= 0);
        const int want = min((int)refill_min, __popcll(active));
        int st = (inst >= 0 && L.t >= L.end) ? BG_OK : -1;  // Done before any step
        for (;;) {
            const uint64_t done = __ballot(inst >= 0 && st >= 0);
            if (__popcll(done) >= want) break;
            if (inst >= 0 && st < 0) st = rk_advance(L, attempts, cap, nacc, nrej);
        }
        if (inst >= 0 && st >= 0) {
            rk_store(L, desc_outputs(dv), (uint64_t)inst, st, nacc, nrej);
            inst = kNeedWork;
        }
    }
}

// Path output for an ensemble: the same persistent refill scheme as
// rk_ensemble_kernel, and each accepted (t, y) the reference's IVPIterator
// would yield (rk.rs:419, ivp.rs:220-238) is written to a per-instance ring
// buffer: every `every`-th accepted step, or only landings on the shared stop
// times and the end.  Stops: the end clamp of rk.rs:366-368 aims at the next
// stop time and an accepted clamped step lands exactly on it.
template <class RHS, class Tab>
__global__ void __launch_bounds__(256)
    rk_path_ensemble_kernel(const bg_ensemble_desc *dv, const bg_path_desc *__restrict__ pd,
                            uint32_t *__restrict__ counter) {
    constexpr int D = RHS::D;
    r8_init();
    RkLane<RHS, Tab> L;
    const int lane = __lane_id();
    const uint64_t max_att = dv->max_attempts;
    const uint32_t cap = (uint32_t)((max_att && max_att < kDefaultMaxAttempts) ? max_att
                                                                                : kDefaultMaxAttempts);
    const uint32_t every = pd->every ? pd->every : 1u;
    const uint32_t capacity = pd->capacity;
    const bool stops_only = (pd->flags & BG_PATH_STOPS_ONLY) != 0;
    const bool inst_major = (pd->flags & BG_PATH_INSTANCE_MAJOR) != 0;
    // global-memory views: FLAT stores would make the next LDS read (the
    // root8 table) wait until they complete
    const __attribute__((address_space(1))) double *stops = gptr(pd->stops);
    __attribute__((address_space(1))) double *path_t = gptr(pd->path_t);
    __attribute__((address_space(1))) double *path_y = gptr(pd->path_y);
    __attribute__((address_space(1))) uint32_t *path_len = gptr(pd->path_len);
    const uint32_t n_stops_all = pd->n_stops;
    int64_t inst = kNeedWork;
    uint32_t attempts = 0, nacc = 0, nrej = 0, next_stop = 0, n_stops = 0, nrec = 0;
    uint32_t every_left = every;  // accepted steps until the next record
    uint32_t slot = 0;            // ring entry of the next record (nrec % capacity)
    double target = 0.0;          // the next stop, else the end
    for (;;) {
        const uint64_t need = __ballot(inst == kNeedWork);
        if (need != 0) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (inst == kNeedWork) {
                const uint64_t pos =
                    (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
                const bg_ensemble_desc di = desc_inputs(dv);
                if (pos < di.n) {
                    inst = di.order ? (int64_t)di.order[pos] : (int64_t)pos;
                    rk_load(L, di, (uint64_t)inst);
                    attempts = nacc = nrej = next_stop = nrec = 0;
                    every_left = every;
                    slot = 0;
                    n_stops = 0;
                    while (n_stops < n_stops_all && stops[n_stops] < L.end) ++n_stops;
                    while (next_stop < n_stops && stops[next_stop] <= L.t) ++next_stop;
                    target = next_stop < n_stops ? stops[next_stop] : L.end;
                } else {
                    inst = kExhausted;
                }
            }
        }
        if (__ballot(inst >= 0) == 0) break;
        int st = (inst >= 0 && L.t >= L.end) ? BG_OK : -1;
        for (;;) {
            if (__ballot(inst >= 0 && st >= 0) != 0) break;
            if (inst >= 0 && st < 0) {
                if (attempts >= cap) {
                    st = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                } else {
                    const bool to_stop = next_stop < n_stops;
                    bool clamped;
                    const int r = L.attempt_to(target, to_stop, clamped);
                    ++attempts;
                    if (r == STEP_ACCEPTED) {
                        ++nacc;
                        bool at_stop = false;
                        if (clamped && to_stop) {
                            L.t = target;
                            ++next_stop;
                            target = next_stop < n_stops ? stops[next_stop] : L.end;
                            at_stop = true;
                        }
                        const bool last = L.t >= L.end;
                        --every_left;
                        const bool rec = stops_only ? (at_stop || last) : (every_left == 0);
                        if (every_left == 0) every_left = every;
                        if (rec) {
                            const uint64_t n = desc_at_use(dv)->n, i = (uint64_t)inst;
                            if (capacity) {
                                const uint64_t e = slot;
                                if (inst_major) {
                                    const uint64_t r = i * capacity + e;
                                    path_t[r] = L.t;
#pragma unroll
                                    for (int c = 0; c < D; ++c) path_y[r * D + c] = L.y[c];
                                } else {
                                    path_t[e * n + i] = L.t;
#pragma unroll
                                    for (int c = 0; c < D; ++c) path_y[(e * D + c) * n + i] = L.y[c];
                                }
                                slot = slot + 1 == capacity ? 0u : slot + 1;
                            }
                            ++nrec;
                        }
                        if (last) st = BG_OK;
                    } else if (r == STEP_REJECTED) {
                        ++nrej;
                    } else {
                        st = r;
                    }
                }
            }
        }
        if (inst >= 0 && st >= 0) {
            rk_store(L, desc_outputs(dv), (uint64_t)inst, st, nacc, nrej);
            if (path_len) path_len[inst] = nrec;
            inst = kNeedWork;
        }
    }
}

// Plain path: one instance per launch, one lane.
template <class RHS, class Tab>
__global__ void __launch_bounds__(64) rk_plain_kernel(const bg_ensemble_desc d, uint64_t k) {
    r8_init();
    if (threadIdx.x != 0) return;
    RkLane<RHS, Tab> L;
    const uint64_t i = d.order ? (uint64_t)d.order[k] : k;
    const uint32_t cap =
        (uint32_t)((d.max_attempts && d.max_attempts < kDefaultMaxAttempts) ? d.max_attempts
                                                                           : kDefaultMaxAttempts);
    rk_load(L, d, i);
    uint32_t attempts = 0;
    uint32_t nacc = 0, nrej = 0;
    int st = L.t >= L.end ? BG_OK : -1;
    while (st < 0) st = rk_advance(L, attempts, cap, nacc, nrej);
    rk_store(L, d, i, st, nacc, nrej);
}

// Single-instance path production: yields every accepted (t, y) into the
// path buffer; stops when the buffer is full (resumable) or the solve ends.
template <class RHS, class Tab>
__global__ void __launch_bounds__(64)
    rk_path_kernel(PathState *__restrict__ st, double *__restrict__ path_t,
                   double *__restrict__ path_y, uint32_t cap) {
    constexpr int D = RHS::D, S = Tab::S;
    r8_init();
    if (threadIdx.x != 0) return;
    RkLane<RHS, Tab> L;
    L.t = st->t;
    L.dt = st->dt;
    L.end = st->end;
    L.tol = st->tol;
    L.dt_min = st->dt_min;
    L.dt_max = st->dt_max;
#pragma unroll
    for (int k = 0; k < D; ++k) L.y[k] = st->y[k];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < D; ++k) L.K[s][k] = st->K[s][k];
#pragma unroll
    for (int k = 0; k < RHS::P; ++k) L.p[k] = st->p[k];
    uint32_t attempts = (uint32_t)st->attempts;
    const uint32_t max_att = (uint32_t)st->max_attempts;
    uint32_t nacc = st->n_accepted, nrej = st->n_rejected, count = 0;
    int status = BG_OK, finished = st->finished;
    while (!finished && count < cap) {
        if (L.t >= L.end) {
            finished = 1;
            break;
        }
        const uint32_t before = nacc;
        const int r = rk_advance(L, attempts, max_att, nacc, nrej);
        if (nacc != before) {
            path_t[count] = L.t;
#pragma unroll
            for (int k = 0; k < D; ++k) path_y[(uint64_t)count * D + k] = L.y[k];
            ++count;
        }
        if (r > 0) {
            status = r;
            finished = 1;
        } else if (r == BG_OK) {
            finished = 1;
        }
    }
    st->t = L.t;
    st->dt = L.dt;
#pragma unroll
    for (int k = 0; k < D; ++k) st->y[k] = L.y[k];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < D; ++k) st->K[s][k] = L.K[s][k];
    st->attempts = attempts;
    st->n_accepted = nacc;
    st->n_rejected = nrej;
    st->status = status;
    st->finished = finished;
    st->count = count;
}

// Work-queue reset and descriptor copy into the per-launch workspace
// (a kernel rather than memset/memcpy so the launch sequence stays capturable).
__global__ void rk_setup_kernel(uint32_t *counter, bg_ensemble_desc *dst, const bg_ensemble_desc d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *dst = d;
    }
}

// ---------------------------------------------------------------------------
// Dispatch (method, rhs, dim, semantics) -> template instance.
// ---------------------------------------------------------------------------
template <class V>
static int visit_rk(int method, int rhs, int semantics, uint32_t dim, V &&v) {
    if (semantics != BG_AS_BUILT && semantics != BG_INTENDED) return BG_ERR_INVALID_ARGUMENT;
    auto with_tab = [&](auto rtag) -> int {
        if (method == BG_RK45)
            return semantics == BG_AS_BUILT ? v(rtag, Tag<Tableau<Fehlberg45, true>>{})
                                            : v(rtag, Tag<Tableau<Fehlberg45, false>>{});
        if (method == BG_RK23)
            return semantics == BG_AS_BUILT ? v(rtag, Tag<Tableau<BogackiShampine23, true>>{})
                                            : v(rtag, Tag<Tableau<BogackiShampine23, false>>{});
        return BG_ERR_UNSUPPORTED;
    };
#define BG_SCALAR_FAMILY(ID, TPL)                          \
    case ID:                                               \
        if (dim == 1) return with_tab(Tag<TPL<1>>{});      \
        if (dim == 2) return with_tab(Tag<TPL<2>>{});      \
        return BG_ERR_UNSUPPORTED;
    switch (rhs) {
        BG_SCALAR_FAMILY(BG_RHS_EXP, RhsExp)
        BG_SCALAR_FAMILY(BG_RHS_NEG, RhsNeg)
        BG_SCALAR_FAMILY(BG_RHS_QUADRATIC, RhsQuadratic)
        BG_SCALAR_FAMILY(BG_RHS_COS, RhsCos)
        BG_SCALAR_FAMILY(BG_RHS_FAIL_AFTER, RhsFailAfter)
    case BG_RHS_LORENZ:
        return dim == 3 ? with_tab(Tag<RhsLorenz>{}) : BG_ERR_INVALID_ARGUMENT;
    case BG_RHS_LOTKA_VOLTERRA:
        return dim == 2 ? with_tab(Tag<RhsLotkaVolterra>{}) : BG_ERR_INVALID_ARGUMENT;
    case BG_RHS_OSCILLATOR:
        return dim == 2 ? with_tab(Tag<RhsOscillator>{}) : BG_ERR_INVALID_ARGUMENT;
    case BG_RHS_ROBERTSON:
        return dim == 3 ? with_tab(Tag<RhsRobertson>{}) : BG_ERR_INVALID_ARGUMENT;
    case BG_RHS_TWO_LEVEL:
        return dim == 4 ? with_tab(Tag<RhsTwoLevel>{}) : BG_ERR_INVALID_ARGUMENT;
    default:
        return BG_ERR_UNSUPPORTED;
    }
#undef BG_SCALAR_FAMILY
}

bool rk_supported(int method, int rhs, uint32_t dim) {
    return visit_rk(method, rhs, BG_INTENDED, dim, [](auto, auto) { return BG_OK; }) == BG_OK;
}

static int check_desc(const bg_ensemble_desc &d, int P) {
    if (d.n == 0) return BG_OK;
    if (d.n > 0x7fffffffull) return BG_ERR_INVALID_ARGUMENT;
    if (!d.y0) return BG_ERR_INVALID_ARGUMENT;
    if (P > 0 && (!d.params || d.n_params < (uint32_t)P)) return BG_ERR_INVALID_ARGUMENT;
    return BG_OK;
}

int rk_launch_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                       const bg_launch &lc, hipStream_t stream, void *ws_in) {
    // the shared-matrix linear right-hand side runs as a tiled MFMA GEMM kernel
    if (rhs == BG_RHS_LINEAR) return linear_launch_ensemble(method, semantics, d, stream, ws_in);
    return visit_rk(method, rhs, semantics, d.dim, [&](auto rtag, auto ttag) -> int {
        using R = typename decltype(rtag)::type;
        using T = typename decltype(ttag)::type;
        int rc = check_desc(d, R::P);
        if (rc != BG_OK || d.n == 0) return rc;
        const uint32_t block = lc.block ? lc.block : 256;
        if (block % 64 || block > 256) return BG_ERR_INVALID_ARGUMENT;
        int dev = 0;
        BG_HIP_CHECK(hipGetDevice(&dev));
        int per_cu = 0;
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, rk_ensemble_kernel<R, T>, (int)block, 0));
        if (lc.blocks_per_cu && (int)lc.blocks_per_cu < per_cu) per_cu = (int)lc.blocks_per_cu;
        if (per_cu < 1) per_cu = 1;
        uint64_t grid = (uint64_t)per_cu * (uint64_t)device_cu_count(dev);
        const uint64_t need = (d.n + block - 1) / block;
        if (grid > need) grid = need;
        // workspace: [0,256) work counter, [256, ...) descriptor copy
        char *ws = (char *)ws_in;
        if (!ws_in) BG_HIP_CHECK(hipMallocAsync((void **)&ws, kLaunchWorkspaceBytes, stream));
        uint32_t *counter = (uint32_t *)ws;
        bg_ensemble_desc *ddesc = (bg_ensemble_desc *)(ws + 256);
        rk_setup_kernel<<<1, 64, 0, stream>>>(counter, ddesc, d);
        const uint32_t refill_min = lc.refill_min ? lc.refill_min : 1u;
        rk_ensemble_kernel<R, T>
            <<<(uint32_t)grid, block, 0, stream>>>(ddesc, counter, refill_min, lc.prio_cutoff);
        const hipError_t launched = hipGetLastError();
        if (!ws_in) BG_HIP_CHECK(hipFreeAsync(ws, stream));
        BG_HIP_CHECK(launched);
        return BG_OK;
    });
}

int rk_launch_plain(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                    hipStream_t stream) {
    return visit_rk(method, rhs, semantics, d.dim, [&](auto rtag, auto ttag) -> int {
        using R = typename decltype(rtag)::type;
        using T = typename decltype(ttag)::type;
        int rc = check_desc(d, R::P);
        if (rc != BG_OK) return rc;
        for (uint64_t k = 0; k < d.n; ++k) {
            rk_plain_kernel<R, T><<<1, 64, 0, stream>>>(d, k);
            BG_HIP_CHECK(hipGetLastError());
            BG_HIP_CHECK(hipStreamSynchronize(stream));
        }
        return BG_OK;
    });
}

int rk_launch_path(int method, int rhs, int semantics, uint32_t dim, PathState *state,
                   double *path_t, double *path_y, uint32_t cap, hipStream_t stream) {
    return visit_rk(method, rhs, semantics, dim, [&](auto rtag, auto ttag) -> int {
        using R = typename decltype(rtag)::type;
        using T = typename decltype(ttag)::type;
        rk_path_kernel<R, T><<<1, 64, 0, stream>>>(state, path_t, path_y, cap);
        BG_HIP_CHECK(hipGetLastError());
        return BG_OK;
    });
}

__global__ void rk_path_setup_kernel(uint32_t *counter, bg_ensemble_desc *dst,
                                     const bg_ensemble_desc d, bg_path_desc *pdst,
                                     const bg_path_desc p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *dst = d;
        *pdst = p;
    }
}

int rk_launch_path_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                            const bg_path_desc &p, hipStream_t stream) {
    return visit_rk(method, rhs, semantics, d.dim, [&](auto rtag, auto ttag) -> int {
        using R = typename decltype(rtag)::type;
        using T = typename decltype(ttag)::type;
        int rc = check_desc(d, R::P);
        if (rc != BG_OK || d.n == 0) return rc;
        if (p.capacity && (!p.path_t || !p.path_y)) return BG_ERR_INVALID_ARGUMENT;
        if (p.n_stops && !p.stops) return BG_ERR_INVALID_ARGUMENT;
        int dev = 0;
        BG_HIP_CHECK(hipGetDevice(&dev));
        int per_cu = 0;
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, rk_path_ensemble_kernel<R, T>, 256, 0));
        if (per_cu < 1) per_cu = 1;
        uint64_t grid = (uint64_t)per_cu * (uint64_t)device_cu_count(dev);
        const uint64_t need = (d.n + 255) / 256;
        if (grid > need) grid = need;
        // workspace: [0,256) counter, then the two descriptor copies
        const size_t off_p = (256 + sizeof(bg_ensemble_desc) + 255) & ~(size_t)255;
        char *ws = nullptr;
        BG_HIP_CHECK(hipMallocAsync((void **)&ws, off_p + sizeof(bg_path_desc), stream));
        uint32_t *counter = (uint32_t *)ws;
        bg_ensemble_desc *ddesc = (bg_ensemble_desc *)(ws + 256);
        bg_path_desc *pdesc = (bg_path_desc *)(ws + off_p);
        rk_path_setup_kernel<<<1, 64, 0, stream>>>(counter, ddesc, d, pdesc, p);
        rk_path_ensemble_kernel<R, T><<<(uint32_t)grid, 256, 0, stream>>>(ddesc, pdesc, counter);
        const hipError_t launched = hipGetLastError();
        BG_HIP_CHECK(hipFreeAsync(ws, stream));
        BG_HIP_CHECK(launched);
        return BG_OK;
    });
}

}  // namespace bg
