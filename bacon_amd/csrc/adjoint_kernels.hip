// adjoint_kernels.hip — discrete adjoint of the batched RK45 step sequence.
//
// Extension without a counterpart in bacon-sci (its ivp module has no
// sensitivity API).  For each instance the loss is
//     L = sum_k |y(t_k) - target_k|^2
// over observation times t_k.  The step controller is not differentiated: the
// gradient is that of the frozen accepted-step map y_{n+1} = Phi(y_n; dt_n, p)
// (bacon's RK step, src/ivp/rk.rs:370-398, with the intended tableau).
//
// Two persistent kernels, one lane per instance, refilled from work counters
// like the ensemble kernel:
//  * forward: the adaptive solve, with the end clamp of rk.rs:366-368 aimed at
//    the next observation time (the state lands exactly on it).  Every K
//    accepted steps it writes a checkpoint (t, dt, next stop, y) to HBM, and
//    it records the size of every accepted step; at each observation it
//    stores the residual y(t_k) - target_k.
//  * backward: segments newest first.  A segment re-runs its K accepted steps
//    from its checkpoint with the recorded step sizes (bit-identical to the
//    forward: the same stage and update operations; the rejected attempts
//    and the controller are not repeated), recording (t, dt, stop flag, y)
//    per step on a per-lane tape in LDS, then sweeps them in reverse:
//    lambda += 2 r_k at an observation, and the stage-by-stage VJP of the
//    step map with the right-hand side's hand-written Jacobian-vector
//    products.
// Storage layouts are structure-of-arrays with the instance (or lane slot) as
// the fastest index, so each lane's loads and stores coalesce with its wave.
#include <rocprim/device/device_radix_sort.hpp>

#include "bg_device.h"
#include "bg_internal.h"

namespace bg {

static constexpr uint32_t kAdjMaxAttempts = 10000000u;

struct AdjWs {
    double *ck;      // [n][maxck][F], F = 3 + D: t, dt, next_stop, y[D]
    double *hs;      // [n][maxck][K] size of each accepted step, negated for a
                     // step that lands on an observation time
    uint32_t *nck;   // [n] checkpoints written
    uint32_t *nobs;  // [n] observations reached
    double *res;     // [n][n_obs][D] residuals y(t_k) - target_k
    double *seg;     // [K][2 + D][slots] backward tape, when not in LDS
    uint64_t slots;
    uint32_t *counters;  // [0] forward queue, [1] backward queue
    const int32_t *order_b;  // [n] backward schedule: most checkpoints first
};

// Backward schedule keys: the forward's checkpoint count per instance is the
// exact number of segments the backward pass will recompute and reverse.
__global__ void adj_order_keys_kernel(const uint32_t *__restrict__ nck, uint64_t n,
                                      uint32_t *__restrict__ keys, int32_t *__restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = nck[i];
    idx[i] = (int32_t)i;
}

// Observation times are read at every instance start (count_stops) and at
// every stop.  With OBS_LDS (n_obs <= kObsLds) they are staged in LDS at
// kernel entry: the reads then neither wait on L2 (a serial scan of 20
// global loads per instance start cost microseconds) nor, being known LDS
// accesses, wait for the loop's outstanding checkpoint stores.  Every thread
// of the block calls it.
static constexpr uint32_t kObsLds = 1024;
template <bool OBS_LDS>
__device__ __forceinline__ const double *stage_obs(double *lds, const double *obs,
                                                   uint32_t n_obs) {
    if constexpr (!OBS_LDS) {
        return obs;
    } else {
        for (uint32_t k = threadIdx.x; k < n_obs; k += blockDim.x) lds[k] = gptr(obs)[k];
        __syncthreads();
        return lds;
    }
}

// Number of observation times strictly before `end` (they become stops):
// a binary search over the increasing times, ~log2(n_obs) dependent reads
// at each instance start instead of one per time before the end.
__device__ __forceinline__ uint32_t count_stops(const double *obs, uint32_t n_obs, double end) {
    uint32_t lo = 0, hi = n_obs;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (obs[mid] < end)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

template <class RHS, class Tab, bool OBS_LDS>
__global__ void __launch_bounds__(256)
    rk_adjoint_forward_kernel(const bg_adjoint_desc *__restrict__ ad, const AdjWs ws) {
    constexpr int D = RHS::D, F = 3 + D;
    r8_init();
    const bg_ensemble_desc &d = ad->base;
    const int lane = __lane_id();
    const uint64_t n = d.n;
    const uint32_t K = ad->checkpoint_every ? ad->checkpoint_every : 16u;
    const uint32_t maxck = ad->max_checkpoints ? ad->max_checkpoints : 128u;
    const uint32_t n_obs = ad->n_obs;
    __shared__ double obs_lds[OBS_LDS ? kObsLds : 1];
    const double *obs = stage_obs<OBS_LDS>(obs_lds, ad->obs_t, n_obs);
    RkLane<RHS, Tab> L;
    int64_t inst = -1;
    uint32_t attempts = 0, nacc = 0, nrej = 0, next_stop = 0, n_stops = 0, nck = 0;
    uint32_t ck_left = K;  // accepted steps until the next checkpoint
    double target = 0.0;  // the next stop, else the end
    for (;;) {
        const uint64_t need = __ballot(inst == -1);
        if (need != 0) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&ws.counters[0], (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (inst == -1) {
                const uint64_t k = (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
                if (k < n) {
                    inst = d.order ? (int64_t)gptr(d.order)[k] : (int64_t)k;
                    const uint64_t i = (uint64_t)inst;
#pragma unroll
                    for (int c = 0; c < D; ++c) L.y[c] = gptr(d.y0)[c * n + i];
#pragma unroll
                    for (int c = 0; c < RHS::P; ++c) L.p[c] = gptr(d.params)[c * n + i];
                    L.t = d.t0 ? gptr(d.t0)[i] : d.t0_s;
                    L.end = d.t_end ? gptr(d.t_end)[i] : d.t_end_s;
                    L.tol = d.tol ? gptr(d.tol)[i] : d.tol_s;
                    L.dt_min = d.dt_min ? gptr(d.dt_min)[i] : d.dt_min_s;
                    L.dt_max = d.dt_max ? gptr(d.dt_max)[i] : d.dt_max_s;
                    L.dt = (L.dt_max + L.dt_min) * 0.5;
                    attempts = nacc = nrej = next_stop = 0;
                    ck_left = K;
                    n_stops = count_stops(obs, n_obs, L.end);
                    target = n_stops > 0 ? obs[0] : L.end;
                    // checkpoint 0: the initial state
                    double *ck0 = ws.ck + i * maxck * F;
                    ck0[0] = L.t;
                    ck0[1] = L.dt;
                    ck0[2] = 0.0;
#pragma unroll
                    for (int c = 0; c < D; ++c) ck0[3 + c] = L.y[c];
                    nck = 1;
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
                if (L.t >= L.end) {
                    st = BG_OK;
                } else if (attempts >= kAdjMaxAttempts) {
                    st = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                } else {
                    const bool to_stop = next_stop < n_stops;
                    bool clamped;
                    double h;
                    const int r = L.attempt_to(target, to_stop, clamped, h);
                    ++attempts;
                    if (r == STEP_ACCEPTED) {
                        const uint64_t i = (uint64_t)inst;
                        // accepted step nacc lies in segment nacc / K < maxck
                        // (the capacity test below ends the solve at the
                        // segment boundary past the last checkpoint)
                        ws.hs[i * maxck * K + nacc] = (clamped && to_stop) ? -h : h;
                        ++nacc;
                        if (clamped && to_stop) {
                            L.t = target;
#pragma unroll
                            for (int c = 0; c < D; ++c)
                                ws.res[(i * n_obs + next_stop) * D + c] = L.y[c];
                            ++next_stop;
                            target = next_stop < n_stops ? obs[next_stop] : L.end;
                        }
                        --ck_left;
                        if (L.t >= L.end) {
                            st = BG_OK;
                        } else if (ck_left == 0) {
                            ck_left = K;
                            if (nck >= maxck) {
                                st = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                            } else {
                                double *ckp = ws.ck + (i * maxck + nck) * F;
                                ckp[0] = L.t;
                                ckp[1] = L.dt;
                                ckp[2] = (double)next_stop;
#pragma unroll
                                for (int c = 0; c < D; ++c) ckp[3 + c] = L.y[c];
                                ++nck;
                            }
                        }
                    } else if (r == STEP_REJECTED) {
                        ++nrej;
                    } else {
                        st = r;
                    }
                }
            }
        }
        if (inst >= 0 && st >= 0) {
            const uint64_t i = (uint64_t)inst;
            // an observation at the end time (after the last stop) scores the
            // final state
            uint32_t reached = next_stop;
            if (st == BG_OK && n_stops < n_obs && obs[n_stops] == L.end) {
#pragma unroll
                for (int c = 0; c < D; ++c) ws.res[(i * n_obs + n_stops) * D + c] = L.y[c];
                reached = n_stops + 1;
            }
            ws.nck[i] = nck;
            ws.nobs[i] = reached;
            if (d.t_out) gptr(d.t_out)[i] = L.t;
            if (d.y_out) {
#pragma unroll
                for (int c = 0; c < D; ++c) gptr(d.y_out)[c * n + i] = L.y[c];
            }
            if (d.status) gptr(d.status)[i] = st;
            if (d.counts) {
                bg_counts cc;
                cc.n_accepted = nacc;
                cc.n_rejected = nrej;
                d.counts[i] = cc;
            }
            inst = -1;
        }
    }
}

// Residuals and loss from the states the forward pass stored at the
// observations it reached (ws.res holds y(t_k) there): res = y - target in
// place, loss = sum over k, c of res^2 in the forward's order, obs_y = y.
// Kept out of the step loop: a target load at one lane's stop would stall
// its whole wave, and in a wave some lane is at a stop on most attempts.
template <int D>
__global__ void adj_residual_kernel(const bg_adjoint_desc *__restrict__ ad, const AdjWs ws) {
    const uint64_t n = ad->base.n;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t n_obs = ad->n_obs, reached = ws.nobs[i];
    double loss = 0.0;
    for (uint32_t k = 0; k < reached; ++k) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const uint64_t o = ((uint64_t)k * D + c) * n + i;
            double *r = ws.res + (i * n_obs + k) * D + c;
            const double y = *r, rr = y - gptr(ad->targets)[o];
            *r = rr;
            loss = fma(rr, rr, loss);
            if (ad->obs_y) gptr(ad->obs_y)[o] = y;
        }
    }
    if (ad->loss) gptr(ad->loss)[i] = loss;
}

// The same, with each block's run of the instance-major residuals
// (kResTile instances x n_obs x D doubles, contiguous) moved between HBM and
// LDS by coalesced loads and stores, so the per-instance loop reads and
// writes LDS instead of one cache line per lane per value.  Identical
// operations, so identical values.
static constexpr uint32_t kResTile = 128;
template <int D>
__global__ void __launch_bounds__(kResTile)
    adj_residual_tiled_kernel(const bg_adjoint_desc *__restrict__ ad, const AdjWs ws) {
    extern __shared__ double tile[];
    const uint64_t n = ad->base.n;
    const uint32_t n_obs = ad->n_obs, per = n_obs * D;
    const uint64_t i0 = (uint64_t)blockIdx.x * kResTile;
    const uint32_t ni = (uint32_t)(n - i0 < kResTile ? n - i0 : kResTile);
    const uint64_t base = i0 * per;
    const uint32_t total = ni * per;
    for (uint32_t e = threadIdx.x; e < total; e += kResTile) tile[e] = ws.res[base + e];
    __syncthreads();
    if (threadIdx.x < ni) {
        const uint64_t i = i0 + threadIdx.x;
        const uint32_t reached = ws.nobs[i];
        double *const mine = tile + (uint64_t)threadIdx.x * per;
        double loss = 0.0;
        for (uint32_t k = 0; k < reached; ++k) {
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const uint64_t o = ((uint64_t)k * D + c) * n + i;
                double *r = mine + k * D + c;
                const double y = *r, rr = y - gptr(ad->targets)[o];
                *r = rr;
                loss = fma(rr, rr, loss);
                if (ad->obs_y) gptr(ad->obs_y)[o] = y;
            }
        }
        if (ad->loss) gptr(ad->loss)[i] = loss;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < total; e += kResTile) ws.res[base + e] = tile[e];
}

// Reverse of one accepted step of the explicit map
//   Y_i = y + sum_{j<i} a_ij K_j,  K_i = dt f(Y_i),  y' = y + sum_i b_i K_i
// given lam = dL/dy'; returns dL/dy in lam and accumulates dL/dp into pb.
template <class RHS, class Tab>
__device__ __forceinline__ void vjp_step(double t, double dt, const double (&y)[RHS::D],
                                         const double *p, double (&lam)[RHS::D], double *pb) {
    constexpr int D = RHS::D, S = Tab::S;
    double Y[S][D], K[S][D], Kb[S][D];
#pragma unroll
    for (int i = 0; i < S; ++i) {
#pragma unroll
        for (int c = 0; c < D; ++c) Y[i][c] = y[c];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const double m = Tab::m(i, j);
            if (m != 0.0) {
#pragma unroll
                for (int c = 0; c < D; ++c) Y[i][c] = fma(m, K[j][c], Y[i][c]);
            }
        }
        double f[D];
        RHS::eval(fma(Tab::c(i), dt, t), Y[i], p, f);
#pragma unroll
        for (int c = 0; c < D; ++c) K[i][c] = f[c] * dt;
    }
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int c = 0; c < D; ++c) Kb[i][c] = Tab::b(i) * lam[c];
    double ysum[D];
#pragma unroll
    for (int c = 0; c < D; ++c) ysum[c] = 0.0;
#pragma unroll
    for (int i = S - 1; i >= 0; --i) {
        double fb[D], yb[D];
#pragma unroll
        for (int c = 0; c < D; ++c) fb[c] = Kb[i][c] * dt;
        RHS::vjp(fma(Tab::c(i), dt, t), Y[i], p, fb, yb, pb);
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const double m = Tab::m(i, j);
            if (m != 0.0) {
#pragma unroll
                for (int c = 0; c < D; ++c) Kb[j][c] = fma(m, yb[c], Kb[j][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < D; ++c) ysum[c] += yb[c];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) lam[c] += ysum[c];
}

// The backward pass's tape: per lane, the K accepted steps of the segment
// being reversed, FT = 2 + D doubles each: t, dt (negated when the step
// landed on an observation time; dt > 0 or +0, so the sign bit is free and
// negation is exact), y.  LDS_TAPE: in LDS, [K][FT][blockDim] (one block of
// 256 lanes holds 128 KiB for Lotka-Volterra at K = 16), else in HBM with
// the lane slot as the fastest index.
template <class RHS, class Tab, bool LDS_TAPE, bool OBS_LDS>
__global__ void __launch_bounds__(256)
    rk_adjoint_backward_kernel(const bg_adjoint_desc *__restrict__ ad, const AdjWs ws) {
    constexpr int D = RHS::D, F = 3 + D, FT = 2 + D, P = RHS::P > 0 ? RHS::P : 1;
    extern __shared__ double tape_lds[];
    r8_init();
    const bg_ensemble_desc &d = ad->base;
    const int lane = __lane_id();
    const uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = d.n;
    double *const tape = LDS_TAPE ? tape_lds + threadIdx.x : ws.seg + slot;
    const uint64_t S = LDS_TAPE ? (uint64_t)blockDim.x : ws.slots;
    const uint32_t K = ad->checkpoint_every ? ad->checkpoint_every : 16u;
    const uint32_t n_obs = ad->n_obs;
    const uint64_t maxck = ad->max_checkpoints ? ad->max_checkpoints : 128u;
    __shared__ double obs_lds[OBS_LDS ? kObsLds : 1];
    const double *obs = stage_obs<OBS_LDS>(obs_lds, ad->obs_t, n_obs);
    RkLane<RHS, Tab, true> L;  // coefficients in SGPR pairs for the replay
    double lam[D], pb[P];
    int64_t inst = -1;
    int seg = -1;
    uint32_t n_stops = 0;  // observation times before the instance's end
    // checkpoint of segment `seg` (t, dt, next stop, y), loaded one segment
    // ahead so the loads overlap the previous segment's work
    double ck[F];
#pragma unroll
    for (int f = 0; f < F; ++f) ck[f] = 0.0;
    auto load_ck = [&](int sg, uint64_t i) {
#pragma unroll
        for (int f = 0; f < F; ++f) ck[f] = ws.ck[(i * maxck + (uint64_t)sg) * F + f];
    };
    // the first four recorded step sizes of segment `sg`, loaded before they
    // are needed (at the instance start, and before the previous segment's
    // reverse sweep); the replay keeps four loads in flight
    double h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;
    auto load_h = [&](int sg, uint64_t i) {
        const double *hs = ws.hs + (i * maxck + (uint64_t)sg) * K;
        h0 = hs[0];
        h1 = K > 1 ? hs[1] : 0.0;
        h2 = K > 2 ? hs[2] : 0.0;
        h3 = K > 3 ? hs[3] : 0.0;
    };
    // residual of the stop the reverse sweep meets next (stops are met in
    // decreasing order, n_stops - 1 first), loaded one stop ahead
    int res_k = -1;
    double rn[D];
#pragma unroll
    for (int c = 0; c < D; ++c) rn[c] = 0.0;
    auto load_res = [&](int k, uint64_t i) {
        res_k = k;
        if (k >= 0) {
#pragma unroll
            for (int c = 0; c < D; ++c) rn[c] = ws.res[(i * n_obs + (uint64_t)k) * D + c];
        }
    };
    for (;;) {
        const uint64_t need = __ballot(inst == -1);
        if (need != 0) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&ws.counters[1], (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (inst == -1) {
                const uint64_t k = (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
                if (k < n) {
                    inst = (int64_t)ws.order_b[k];
                    const uint64_t i = (uint64_t)inst;
#pragma unroll
                    for (int c = 0; c < RHS::P; ++c) L.p[c] = gptr(d.params)[c * n + i];
                    L.end = d.t_end ? gptr(d.t_end)[i] : d.t_end_s;
                    L.tol = d.tol ? gptr(d.tol)[i] : d.tol_s;
                    L.dt_min = d.dt_min ? gptr(d.dt_min)[i] : d.dt_min_s;
                    L.dt_max = d.dt_max ? gptr(d.dt_max)[i] : d.dt_max_s;
#pragma unroll
                    for (int c = 0; c < P; ++c) pb[c] = 0.0;
                    n_stops = count_stops(obs, n_obs, L.end);
                    const bool final_obs = n_stops < n_obs && obs[n_stops] == L.end;
#pragma unroll
                    for (int c = 0; c < D; ++c)
                        lam[c] = final_obs ? 2.0 * ws.res[(i * n_obs + n_stops) * D + c] : 0.0;
                    const int fw = d.status ? gptr(d.status)[i] : BG_OK;
                    seg = fw == BG_OK ? (int)ws.nck[i] - 1 : -1;
                    load_res((int)n_stops - 1, i);
                    if (seg >= 0) {
                        load_ck(seg, i);
                        load_h(seg, i);
                    }
                } else {
                    inst = -2;
                }
            }
        }
        if (__ballot(inst >= 0) == 0) break;
        bool fin = false;
        for (;;) {
            if (__ballot(inst >= 0 && fin) != 0) break;
            if (inst >= 0 && !fin) {
                if (seg < 0) {
                    fin = true;
                } else {
                    const uint64_t i = (uint64_t)inst;
                    // ---- re-run segment `seg` from its checkpoint ----
                    L.t = ck[0];
                    uint32_t next_stop = (uint32_t)ck[2];
#pragma unroll
                    for (int c = 0; c < D; ++c) L.y[c] = ck[3 + c];
                    if (seg > 0) load_ck(seg - 1, i);
                    double target = next_stop < n_stops ? obs[next_stop] : L.end;
                    const double *hs = ws.hs + (i * maxck + (uint64_t)seg) * K;
                    uint32_t cnt = 0;
                    while (cnt < K && L.t < L.end) {
                        const double hk = h0;  // negative: the step lands on a stop
                        h0 = h1;
                        h1 = h2;
                        h2 = h3;
                        h3 = cnt + 4 < K ? hs[cnt + 4] : 0.0;
                        double *rec = tape + ((uint64_t)cnt * FT) * S;
                        rec[0 * S] = L.t;
                        rec[1 * S] = hk;
#pragma unroll
                        for (int c = 0; c < D; ++c) rec[(2 + c) * S] = L.y[c];
                        if (!L.advance(fabs(hk))) break;  // not on a replay of a successful forward
                        if (signbit(hk)) {
                            L.t = target;
                            ++next_stop;
                            target = next_stop < n_stops ? obs[next_stop] : L.end;
                        }
                        ++cnt;
                    }
                    if (seg > 0) load_h(seg - 1, i);
                    // ---- reverse sweep ----
                    // Observation steps take the stop indices in reverse:
                    // the replay ended with next_stop one past the last.
                    // Step k - 1's record is loaded while step k reverses.
                    uint32_t cursor = next_stop;
                    double tk = 0.0, dk = 0.0, yk[D];
#pragma unroll
                    for (int c = 0; c < D; ++c) yk[c] = 0.0;
                    if (cnt > 0) {
                        const double *rec = tape + ((uint64_t)(cnt - 1) * FT) * S;
                        tk = rec[0 * S];
                        dk = rec[1 * S];
#pragma unroll
                        for (int c = 0; c < D; ++c) yk[c] = rec[(2 + c) * S];
                    }
                    for (int k = (int)cnt - 1; k >= 0; --k) {
                        double tn = 0.0, dn = 0.0, yn[D];
#pragma unroll
                        for (int c = 0; c < D; ++c) yn[c] = 0.0;
                        if (k > 0) {
                            const double *rec = tape + ((uint64_t)(k - 1) * FT) * S;
                            tn = rec[0 * S];
                            dn = rec[1 * S];
#pragma unroll
                            for (int c = 0; c < D; ++c) yn[c] = rec[(2 + c) * S];
                        }
                        if (signbit(dk)) {
                            const uint64_t s = --cursor;
                            if ((int)s != res_k) load_res((int)s, i);
#pragma unroll
                            for (int c = 0; c < D; ++c) lam[c] = fma(2.0, rn[c], lam[c]);
                            load_res((int)s - 1, i);
                        }
                        vjp_step<RHS, Tab>(tk, fabs(dk), yk, L.p, lam, pb);
                        tk = tn;
                        dk = dn;
#pragma unroll
                        for (int c = 0; c < D; ++c) yk[c] = yn[c];
                    }
                    --seg;
                    if (seg < 0) fin = true;
                }
            }
        }
        if (inst >= 0 && fin) {
            const uint64_t i = (uint64_t)inst;
            const int fw = d.status ? gptr(d.status)[i] : BG_OK;
            const double nan = __builtin_nan("");
            if (ad->grad_y0) {
#pragma unroll
                for (int c = 0; c < D; ++c) gptr(ad->grad_y0)[c * n + i] = fw == BG_OK ? lam[c] : nan;
            }
            if (ad->grad_params) {
#pragma unroll
                for (int c = 0; c < RHS::P; ++c)
                    gptr(ad->grad_params)[c * n + i] = fw == BG_OK ? pb[c] : nan;
            }
            inst = -1;
        }
    }
}

__global__ void adj_setup_kernel(uint32_t *counters, bg_adjoint_desc *dst, const bg_adjoint_desc d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counters[0] = 0u;
        counters[1] = 0u;
        *dst = d;
    }
}

template <class T>
struct ATag {
    using type = T;
};

int rk_launch_adjoint(int method, int rhs, const bg_adjoint_desc &ad, hipStream_t stream) {
    if (method != BG_RK45) return BG_ERR_UNSUPPORTED;
    const bg_ensemble_desc &d = ad.base;
    auto run = [&](auto rtag) -> int {
        using R = typename decltype(rtag)::type;
        using T = Tableau<Fehlberg45, false>;
        constexpr int D = R::D, F = 3 + D, FT = 2 + D;
        if (d.dim != (uint32_t)D) return BG_ERR_INVALID_ARGUMENT;
        if (d.n == 0) return BG_OK;
        if (d.n > 0x7fffffffull || !d.y0 || !d.params || d.n_params < (uint32_t)R::P ||
            !d.status || (ad.n_obs && (!ad.obs_t || !ad.targets)))
            return BG_ERR_INVALID_ARGUMENT;
        const uint64_t n = d.n;
        const uint64_t K = ad.checkpoint_every ? ad.checkpoint_every : 16u;
        const uint64_t maxck = ad.max_checkpoints ? ad.max_checkpoints : 128u;
        int dev = 0;
        BG_HIP_CHECK(hipGetDevice(&dev));
        const int cus = device_cu_count(dev);
        // The backward tape goes to LDS when a 256-lane block's fits (K = 16:
        // Lotka-Volterra 128 KiB); larger tapes stay in HBM.
        int lds_max = 0;
        BG_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t tape_bytes = (size_t)K * FT * sizeof(double) * 256;
        const bool lds_tape = tape_bytes + sizeof(double) * kObsLds + 1024 <= (size_t)lds_max;
        const bool obs_lds = ad.n_obs <= kObsLds;
        auto fwd = obs_lds ? rk_adjoint_forward_kernel<R, T, true>
                           : rk_adjoint_forward_kernel<R, T, false>;
        auto bwd = lds_tape ? (obs_lds ? rk_adjoint_backward_kernel<R, T, true, true>
                                       : rk_adjoint_backward_kernel<R, T, true, false>)
                            : (obs_lds ? rk_adjoint_backward_kernel<R, T, false, true>
                                       : rk_adjoint_backward_kernel<R, T, false, false>);
        const size_t dyn_lds = lds_tape ? tape_bytes : 0;
        if (lds_tape)
            BG_HIP_CHECK(hipFuncSetAttribute((const void *)bwd,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)dyn_lds));
        int per_cu_f = 0, per_cu_b = 0;
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_f, fwd, 256, 0));
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_b, bwd, 256, dyn_lds));
        per_cu_f = per_cu_f < 1 ? 1 : per_cu_f;
        per_cu_b = per_cu_b < 1 ? 1 : per_cu_b;
        // Fewer resident lanes than the occupancy limit: each lane works
        // through several instances in schedule order, so lanes finish
        // together.  Measured on Lotka-Volterra: forward best at 2 blocks
        // per CU; backward best at 1 (with an HBM tape the tape's L2
        // footprint grows with the resident lanes).
        const int want_f = ad.blocks_per_cu ? (int)ad.blocks_per_cu : 2;
        const int want_b = ad.blocks_per_cu ? (int)ad.blocks_per_cu : 1;
        if (per_cu_f > want_f) per_cu_f = want_f;
        if (per_cu_b > want_b) per_cu_b = want_b;
        uint64_t grid_f = (uint64_t)per_cu_f * cus, grid_b = (uint64_t)per_cu_b * cus;
        const uint64_t need = (n + 255) / 256;
        if (grid_f > need) grid_f = need;
        if (grid_b > need) grid_b = need;
        const uint64_t slots = lds_tape ? 0 : grid_b * 256;
        const uint64_t n_res = (uint64_t)(ad.n_obs ? ad.n_obs : 1) * D * n;
        size_t sort_bytes = 0;
        BG_HIP_CHECK(rocprim::radix_sort_pairs_desc(
            nullptr, sort_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
            (const int32_t *)nullptr, (int32_t *)nullptr, (uint32_t)n, 0, 32, stream));
        const size_t bytes = 256 + sizeof(bg_adjoint_desc) + 64 +
                             sizeof(double) * (maxck * F * n + maxck * K * n + n_res +
                                               K * FT * slots) +
                             sizeof(uint32_t) * n + 256 + 5 * (sizeof(uint32_t) * n + 256) +
                             sort_bytes;
        char *buf = nullptr;
        BG_HIP_CHECK(hipMallocAsync((void **)&buf, bytes, stream));
        AdjWs ws;
        ws.counters = (uint32_t *)buf;
        bg_adjoint_desc *dd = (bg_adjoint_desc *)(buf + 256);
        char *p = buf + 256 + sizeof(bg_adjoint_desc);
        p += (64 - ((uintptr_t)p % 64)) % 64;
        ws.ck = (double *)p;
        ws.hs = ws.ck + maxck * F * n;
        ws.res = ws.hs + maxck * K * n;
        ws.seg = ws.res + n_res;
        ws.nck = (uint32_t *)(ws.seg + K * FT * slots);
        ws.slots = slots;
        auto align = [](char *q) { return q + (256 - ((uintptr_t)q % 256)) % 256; };
        char *q = align((char *)(ws.nck + n));
        uint32_t *keys = (uint32_t *)q;
        q = align(q + sizeof(uint32_t) * n);
        uint32_t *keys_sorted = (uint32_t *)q;
        q = align(q + sizeof(uint32_t) * n);
        int32_t *idx = (int32_t *)q;
        q = align(q + sizeof(uint32_t) * n);
        int32_t *order_b = (int32_t *)q;
        q = align(q + sizeof(uint32_t) * n);
        ws.nobs = (uint32_t *)q;
        q = align(q + sizeof(uint32_t) * n);
        void *sort_tmp = q;
        ws.order_b = order_b;
        adj_setup_kernel<<<1, 64, 0, stream>>>(ws.counters, dd, ad);
        fwd<<<(uint32_t)grid_f, 256, 0, stream>>>(dd, ws);
        // residuals through LDS tiles when a tile fits 64 KiB (n_obs <= 25
        // at D = 2 … kResTile x n_obs x D doubles)
        const size_t res_tile = (size_t)kResTile * (ad.n_obs ? ad.n_obs : 1) * D * sizeof(double);
        if (ad.n_obs && res_tile <= 65536)
            adj_residual_tiled_kernel<D>
                <<<(uint32_t)((n + kResTile - 1) / kResTile), kResTile, res_tile, stream>>>(dd, ws);
        else
            adj_residual_kernel<D><<<(uint32_t)((n + 255) / 256), 256, 0, stream>>>(dd, ws);
        adj_order_keys_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, stream>>>(ws.nck, n, keys, idx);
        hipError_t e = rocprim::radix_sort_pairs_desc(sort_tmp, sort_bytes, keys, keys_sorted,
                                                      idx, order_b, (uint32_t)n, 0, 32, stream);
        if (e == hipSuccess) {
            bwd<<<(uint32_t)grid_b, 256, dyn_lds, stream>>>(dd, ws);
            e = hipGetLastError();
        }
        // the workspace goes back on every path
        BG_HIP_CHECK(hipFreeAsync(buf, stream));
        BG_HIP_CHECK(e);
        return BG_OK;
    };
    switch (rhs) {
    case BG_RHS_LOTKA_VOLTERRA:
        return run(ATag<RhsLotkaVolterra>{});
    case BG_RHS_LORENZ:
        return run(ATag<RhsLorenz>{});
    default:
        return BG_ERR_UNSUPPORTED;
    }
}

}  // namespace bg
