// linear_kernels.hip — batched RK45 for y' = A y + b with one shared dense A
// (N = 32, 64 or 128), the stage right-hand side as an FP64 MFMA GEMM.
//
// Hot path: bacon-sci's RungeKuttaSolver::step (src/ivp/rk.rs:361-423) for a
// linear right-hand side shared by every instance.  For a tile of T = 16
// instances the six stage evaluations are products A * Y with Y the N x 16
// matrix of stage inputs: dense FP64 GEMMs on v_mfma_f64_16x16x4f64.
//
// Workgroup = N / 32 waves = one tile of 16 instances (the MFMA N dimension).
//  * Wave w owns rows [32w, 32w+32) of A as MFMA A-operands, resident in
//    registers for the whole launch (2 M-tiles x N/4 k-steps: 64 doubles per
//    lane at N = 128): A is read from HBM once per workgroup.  Beyond N = 128
//    A no longer fits the registers of a workgroup's waves.
//  * v_mfma_f64_16x16x4f64 operand layout on gfx950 (D = A B + C, 16x4 by
//    4x16; checked with a probe kernel): lane l supplies A[l % 16][l / 16]
//    and B[l / 16][l % 16], and register r of lane l holds
//    D[4 r + l / 16][l % 16].
//  * Lane l of a wave holds, for instance column j = l % 16, the 8 rows
//    32w + 16 mt + 4 r + l/16 (mt = 0,1; r = 0..3) of y, of the six stage
//    vectors K_i and of the stage input — exactly the MFMA accumulator layout,
//    so the RK combinations are per-lane FMAs on the GEMM outputs.
//  * Each stage writes its input Y into an N x 16 LDS tile (the B operands
//    of all the waves), barrier, 2 x N/4 MFMAs per wave, barrier.
//  * The error norm of each instance reduces over its 8 rows per lane, over
//    the 4 lanes of the same column (shuffles), and over the waves (LDS) in
//    a fixed order, so every lane holding the column computes the same
//    controller decision.  The controller is RkLane's (squared-norm form).
//  * Instances are taken per column from a global counter: a column whose
//    instance finishes stores it and takes the next, so the 16 columns of a
//    GEMM stay filled until the counter runs out.  A column's arithmetic
//    does not depend on the other columns, so results do not depend on
//    which column or when an instance runs.
#include "bg_device.h"
#include "bg_internal.h"

namespace bg {

typedef double double4_t __attribute__((ext_vector_type(4)));

static constexpr int LT = 16;       // instances per tile
static constexpr int LRW = 32;      // rows per wave (two 16-row M-tiles)
static constexpr uint32_t kLinMaxAttempts = 10000000u;

// LWAVES waves per workgroup, dimension LN = 32 LWAVES.
template <class Tab, int LWAVES>
__global__ void __launch_bounds__(64 * LWAVES)
    linear_rk_kernel(const bg_ensemble_desc *__restrict__ dp, uint32_t *__restrict__ counter) {
    constexpr int S = Tab::S;
    constexpr int LN = LRW * LWAVES;
    const bg_ensemble_desc &d = *dp;
    __shared__ double ytile[LN][LT];
    __shared__ double red[LWAVES][LT];
    __shared__ uint32_t tile_base;

    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const int l = tid & 63;
    const int col = l & 15;   // instance column in the tile
    const int q = l >> 4;     // row group
    const uint64_t n = d.n;
    const double *A = d.shared;           // row-major N x N
    const double *bvec = d.shared + LN * LN;

    // ---- A operands, resident: aop[mt][ks] = A[32w + 16mt + (l%16)][4ks + l/16]
    double aop[2][LN / 4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int ks = 0; ks < LN / 4; ++ks)
            aop[mt][ks] = A[(size_t)(LRW * w + 16 * mt + (l & 15)) * LN + 4 * ks + (l >> 4)];
    // rows owned by this lane and their b entries
    double bl[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) bl[mt][r] = bvec[LRW * w + 16 * mt + 4 * r + q];

    const uint32_t max_att =
        (uint32_t)((d.max_attempts && d.max_attempts < kLinMaxAttempts) ? d.max_attempts
                                                                       : kLinMaxAttempts);
    // Column state.  A column whose instance finishes stores it and takes
    // the next one at the top of the following iteration, so no column
    // rides along idle while the work counter lasts.
    double y[2][4], K[S][2][4];
    double t = 0, dt = 0, end = 0, tol = 1, dt_min = 0, dt_max = 0;
    int status = BG_OK;  // -1 running; >= 0 finished (or no instance)
    bool valid = false;  // the column holds an instance not yet stored
    uint64_t inst = 0;
    uint32_t attempts = 0, nacc = 0, nrej = 0;
    bool exhausted = false;  // block-uniform: the counter has passed n
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) y[mt][r] = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) K[i][mt][r] = 0.0;

    for (;;) {
        // ---- store finished columns
        if (valid && status >= 0) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (d.y_out) d.y_out[(uint64_t)(LRW * w + 16 * mt + 4 * r + q) * n + inst] = y[mt][r];
            if (w == 0 && q == 0) {
                if (d.t_out) d.t_out[inst] = t;
                if (d.status) d.status[inst] = status;
                if (d.counts) {
                    bg_counts c;
                    c.n_accepted = nacc;
                    c.n_rejected = nrej;
                    d.counts[inst] = c;
                }
            }
            valid = false;
        }
        // ---- refill idle columns from the work counter
        if (!exhausted) {
            red[w][col] = status >= 0 ? 1.0 : 0.0;
            __syncthreads();
            uint32_t idle = 0;
#pragma unroll
            for (int c = 0; c < LT; ++c) idle |= (red[0][c] != 0.0 ? 1u : 0u) << c;
            const uint32_t cnt = (uint32_t)__popc(idle);
            if (tid == 0 && cnt) tile_base = atomicAdd(counter, cnt);
            __syncthreads();
            if (cnt) {
                const uint64_t base = tile_base;
                if (base + cnt >= n) exhausted = true;
                if ((idle >> col) & 1u) {
                    const uint64_t pos = base + (uint64_t)__popc(idle & ((1u << col) - 1u));
                    if (pos < n) {
                        valid = true;
                        inst = d.order ? (uint64_t)d.order[pos] : pos;
#pragma unroll
                        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                y[mt][r] = d.y0[(uint64_t)(LRW * w + 16 * mt + 4 * r + q) * n + inst];
                        t = d.t0 ? d.t0[inst] : d.t0_s;
                        end = d.t_end ? d.t_end[inst] : d.t_end_s;
                        tol = d.tol ? d.tol[inst] : d.tol_s;
                        dt_min = d.dt_min ? d.dt_min[inst] : d.dt_min_s;
                        dt_max = d.dt_max ? d.dt_max[inst] : d.dt_max_s;
                        dt = (dt_max + dt_min) * 0.5;  // rk.rs:315
#pragma unroll
                        for (int i = 0; i < S; ++i)
#pragma unroll
                            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                                for (int r = 0; r < 4; ++r) K[i][mt][r] = 0.0;  // rk.rs:323-327
                        attempts = nacc = nrej = 0;
                        status = t >= end ? BG_OK : -1;
                    }
                }
            }
            __syncthreads();
        }
        {
            // any column running?
            red[w][col] = status < 0 ? 1.0 : 0.0;
            __syncthreads();
            double running = 0.0;
#pragma unroll
            for (int c = 0; c < LT; ++c) running += red[0][c];
            __syncthreads();
            if (running == 0.0) {
                if (exhausted) break;
                continue;  // every new instance started at its end: store, refill
            }

            const bool active = status < 0;
            if (active) {
                if (attempts >= max_att)
                    status = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                else if (t + dt >= end)  // rk.rs:366-368
                    dt = end - t;
            }
            const bool step = status < 0;
            // ---- stages (rk.rs:370-384)
#pragma unroll
            for (int i = 0; i < S; ++i) {
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        double v = y[mt][r];
#pragma unroll
                        for (int j = 0; j < S; ++j) {
                            const double m = Tab::m(i, j);
                            if (m != 0.0) v = fma(m, K[j][mt][r], v);
                        }
                        ytile[LRW * w + 16 * mt + 4 * r + q][col] = v;
                    }
                __syncthreads();
                double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < LN / 4; ++ks) {
                    const double bop = ytile[4 * ks + (l >> 4)][l & 15];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[0][ks], bop, acc0, 0, 0, 0);
                }
#pragma unroll
                for (int ks = 0; ks < LN / 4; ++ks) {
                    const double bop = ytile[4 * ks + (l >> 4)][l & 15];
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[1][ks], bop, acc1, 0, 0, 0);
                }
                // the first M-tile's scaling overlaps the second chain's MFMAs
#pragma unroll
                for (int r = 0; r < 4; ++r) K[i][0][r] = (acc0[r] + bl[0][r]) * dt;
                __syncthreads();
#pragma unroll
                for (int r = 0; r < 4; ++r) K[i][1][r] = (acc1[r] + bl[1][r]) * dt;
            }
            // ---- error (rk.rs:386-390), squared, reduced over the column
            constexpr int e0 = Tab::e_first();
            double ss = 0.0;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double E = Tab::e(e0) * K[e0][mt][r];
#pragma unroll
                    for (int j = e0 + 1; j < S; ++j) {
                        const double e = Tab::e(j);
                        if (e != 0.0) E = fma(e, K[j][mt][r], E);
                    }
                    ss = fma(E, E, ss);
                }
            ss += __shfl_xor(ss, 16);
            ss += __shfl_xor(ss, 32);
            if (q == 0) red[w][col] = ss;
            __syncthreads();
            ss = red[0][col];
#pragma unroll
            for (int v = 1; v < LWAVES; ++v) ss += red[v][col];
            __syncthreads();
            // ---- controller (rk.rs:392-416), squared form as RkLane::attempt_to
            if (step) {
                ++attempts;
                const double qq = tol * dt;
                const double q2 = qq * qq;
                const bool accepted = ss <= q2;
                const double rr = q2 / ss;
                const double f0 = rr <= 1e-8 ? 0.1 : (rr >= 65536.0 ? 4.0 : root8(rr));
                const double tang = fma(rr - 1.0, 0.125, 1.0);
                const double f = fmin(f0, tang);
                const double factor = accepted ? fmax(f, 1.0) : fmin(f, 0x1.fffffffffffffp-1);
                if (accepted) {
                    t += dt;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        const double bw = Tab::b(j);
                        if (bw != 0.0) {
#pragma unroll
                            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                                for (int r = 0; r < 4; ++r) y[mt][r] = fma(bw, K[j][mt][r], y[mt][r]);
                        }
                    }
                }
                dt = fmin(dt * factor, dt_max);
                // rk.rs:414-416; an accepted step that then fails the dt_min
                // test is reported as the failure, like the generic kernels
                if (dt < dt_min && t < end) {
                    status = BG_MINIMUM_TIME_DELTA_EXCEEDED;
                } else if (accepted) {
                    ++nacc;
                    if (t >= end) status = BG_OK;  // rk.rs:362-364 on the next call
                } else {
                    ++nrej;
                }
            }
        }
    }
}

template <class T>
struct LinTab {
    using type = T;
};

__global__ void linear_setup_kernel(uint32_t *counter, bg_ensemble_desc *dst,
                                    const bg_ensemble_desc d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *dst = d;
    }
}

bool linear_supported(uint32_t dim) { return dim == 32 || dim == 64 || dim == 128; }

int linear_launch_ensemble(int method, int semantics, const bg_ensemble_desc &d,
                           hipStream_t stream, void *ws_in) {
    if (method != BG_RK45) return BG_ERR_UNSUPPORTED;
    if (!linear_supported(d.dim)) return BG_ERR_UNSUPPORTED;
    if (d.n == 0) return BG_OK;
    if (d.n > 0x7fffffffull || !d.y0 || !d.shared) return BG_ERR_INVALID_ARGUMENT;
    const int threads = 2 * (int)d.dim;  // 64 per 32 rows
    auto run = [&](auto *kern) -> int {
        int dev = 0;
        BG_HIP_CHECK(hipGetDevice(&dev));
        int per_cu = 0;
        BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
        if (per_cu < 1) per_cu = 1;
        uint64_t grid = (uint64_t)per_cu * (uint64_t)device_cu_count(dev);
        const uint64_t tiles = (d.n + LT - 1) / LT;
        if (grid > tiles) grid = tiles;
        char *ws = (char *)ws_in;
        if (!ws_in) BG_HIP_CHECK(hipMallocAsync((void **)&ws, kLaunchWorkspaceBytes, stream));
        uint32_t *counter = (uint32_t *)ws;
        bg_ensemble_desc *ddesc = (bg_ensemble_desc *)(ws + 256);
        linear_setup_kernel<<<1, 64, 0, stream>>>(counter, ddesc, d);
        kern<<<(uint32_t)grid, threads, 0, stream>>>(ddesc, counter);
        const hipError_t launched = hipGetLastError();
        if (!ws_in) BG_HIP_CHECK(hipFreeAsync(ws, stream));
        BG_HIP_CHECK(launched);
        return BG_OK;
    };
    auto by_dim = [&](auto tab) -> int {
        using T = typename decltype(tab)::type;
        if (d.dim == 32) return run(linear_rk_kernel<T, 1>);
        if (d.dim == 64) return run(linear_rk_kernel<T, 2>);
        return run(linear_rk_kernel<T, 4>);
    };
    if (semantics == BG_AS_BUILT) return by_dim(LinTab<Tableau<Fehlberg45, true>>{});
    if (semantics == BG_INTENDED) return by_dim(LinTab<Tableau<Fehlberg45, false>>{});
    return BG_ERR_INVALID_ARGUMENT;
}

}  // namespace bg
