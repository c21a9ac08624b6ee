// tune.hip — the autotuner for batched RK solves: a fitted schedule model
// and a measured launch shape.
//
// Schedule.  A persistent launch hands instances to lanes in schedule order
// and a lane takes the next one when its instance finishes, so the launch ends
// when its longest-running lane does.  Handing out the longest solves first
// (LPT) keeps the lanes' loads even.  The cost of a solve is predicted by
//     log cost = c0 + c1 log(t_end - t0) + c2 log(tol) + sum_k c_{3+k} p_k
// with the coefficients fitted by least squares to the step attempts of a
// solve of the same workload (a deterministic subsample of its instances).
// The per-instance parameters enter because they set the dynamics (for
// Lorenz-63, rho moves the attempt count by ~3 % per unit).
//
// Launch shape.  Each candidate (block size, resident blocks per CU) is timed
// on the caller's whole problem with the fitted schedule; the fastest wins.
// Fewer resident waves than the occupancy limit can win: the VALU is
// saturated by about three waves per SIMD, and fewer lanes mean each lane
// works through more instances, so the end-of-queue tail is shorter.
//
// Results of an RK solve are bit-identical for every schedule and shape (a
// lane's arithmetic does not depend on them), so every trial writes a valid
// solution into the caller's outputs.
//
// The adjoint's tuner (adjoint_autotune) times the gradient solve over
// checkpoint intervals and resident blocks per CU; gradients are the same
// bits for all of them.
//
// BDF solves get the same treatment (bdf_autotune): the cost target is the
// corrector's work, log(16 x Jacobian inverses + Broyden iterations + 1)
// (an inverse costs about sixteen iterations' instructions on the 32-wide
// layout), and the launch candidates are caps on the resident waves per CU
// (one instance per wave, or per lane group for small systems).  A BDF
// group's arithmetic is its own, so every trial is bit-identical too.
#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "bg_internal.h"

namespace bg {

struct CostModel {
    double c[3 + BG_TUNE_MAX_PARAMS];
    uint32_t n_params;
};

__device__ __forceinline__ double desc_at(const double *a, double s, uint64_t i) {
    return a ? a[i] : s;
}

// Features of instance i: [log(t_end - t0), log(tol), p_0 .. p_{P-1}].
__device__ __forceinline__ void instance_features(const bg_ensemble_desc &d, uint64_t i,
                                                  uint32_t P, double *f) {
    const double span = desc_at(d.t_end, d.t_end_s, i) - desc_at(d.t0, d.t0_s, i);
    f[0] = log(span > 0.0 ? span : 1e-300);
    const double tol = desc_at(d.tol, d.tol_s, i);
    f[1] = log(tol > 0.0 ? tol : 1e-300);
    for (uint32_t k = 0; k < P; ++k) f[2 + k] = d.params[(uint64_t)k * d.n + i];
}

// Sample rows for the fit: row j = instance j*stride, columns = features,
// then log(attempts + 1).
__global__ void tune_sample_kernel(const bg_ensemble_desc d, const bg_counts *__restrict__ counts,
                                   uint64_t stride, uint32_t m, uint32_t P,
                                   double *__restrict__ rows) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t i = (uint64_t)j * stride;
    double *r = rows + (uint64_t)j * (P + 3);
    instance_features(d, i, P, r);
    const double att = (double)counts[i].n_accepted + (double)counts[i].n_rejected;
    r[P + 2] = log(att + 1.0);
}

// The same rows from a BDF solve's work counters (.n_accepted = Broyden
// iterations, .n_rejected = Jacobian inverses).
__global__ void tune_sample_bdf_kernel(const bg_ensemble_desc d, const bg_counts *__restrict__ work,
                                       uint64_t stride, uint32_t m, uint32_t P,
                                       double *__restrict__ rows) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t i = (uint64_t)j * stride;
    double *r = rows + (uint64_t)j * (P + 3);
    instance_features(d, i, P, r);
    const double cost = 16.0 * (double)work[i].n_rejected + (double)work[i].n_accepted;
    r[P + 2] = log(cost + 1.0);
}

// Sort keys: predicted log cost (descending sort = longest first).
__global__ void tune_cost_kernel(const bg_ensemble_desc d, const CostModel cm,
                                 float *__restrict__ keys, int32_t *__restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    double f[2 + BG_TUNE_MAX_PARAMS];
    instance_features(d, i, cm.n_params, f);
    double c = cm.c[0] + cm.c[1] * f[0] + cm.c[2] * f[1];
    for (uint32_t k = 0; k < cm.n_params; ++k) c += cm.c[3 + k] * f[2 + k];
    keys[i] = (float)c;
    idx[i] = (int32_t)i;
}

static CostModel model_of(const bg_tuning &tn) {
    CostModel cm;
    memset(&cm, 0, sizeof cm);
    memcpy(cm.c, tn.coef, sizeof cm.c);
    cm.n_params = tn.n_params;
    return cm;
}

// Schedule permutation into `order` (device, [n]).  Workspace from the
// stream-ordered allocator, released on the stream.
static int schedule(const bg_tuning &tn, const bg_ensemble_desc &d, int32_t *order,
                    hipStream_t stream) {
    const uint64_t n = d.n;
    if (n == 0) return BG_OK;
    if (n > 0x7fffffffull || !order) return BG_ERR_INVALID_ARGUMENT;
    if (tn.n_params > BG_TUNE_MAX_PARAMS || (tn.n_params && (!d.params || d.n_params < tn.n_params)))
        return BG_ERR_INVALID_ARGUMENT;
    size_t sort_bytes = 0;
    BG_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, (const float *)nullptr,
                                                (float *)nullptr, (const int32_t *)nullptr,
                                                (int32_t *)nullptr, (uint32_t)n, 0, 32, stream));
    const size_t kbytes = ((n * sizeof(float)) + 255) & ~(size_t)255;
    const size_t ibytes = ((n * sizeof(int32_t)) + 255) & ~(size_t)255;
    char *ws = nullptr;
    BG_HIP_CHECK(hipMallocAsync((void **)&ws, 2 * kbytes + ibytes + sort_bytes, stream));
    float *keys = (float *)ws;
    float *keys_sorted = (float *)(ws + kbytes);
    int32_t *idx = (int32_t *)(ws + 2 * kbytes);
    void *tmp = ws + 2 * kbytes + ibytes;
    tune_cost_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, stream>>>(d, model_of(tn), keys, idx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs_desc(tmp, sort_bytes, keys, keys_sorted, idx, order,
                                           (uint32_t)n, 0, 32, stream);
    // the workspace goes back on every path
    BG_HIP_CHECK(hipFreeAsync(ws, stream));
    BG_HIP_CHECK(e);
    return BG_OK;
}

// Least squares for the cost model over rows [m][P+3] (features, target),
// through the normal equations with a small ridge; features that do not vary
// over the sample get a zero coefficient.
static void fit_model(const std::vector<double> &rows, uint32_t m, uint32_t P, bg_tuning &tn) {
    const uint32_t F = P + 2;  // features without the intercept
    const uint32_t W = P + 3;
    std::vector<int> use;      // columns of the design matrix: -1 = intercept
    use.push_back(-1);
    for (uint32_t c = 0; c < F; ++c) {
        double mn = rows[c], mx = rows[c];
        for (uint32_t j = 1; j < m; ++j) {
            const double v = rows[(size_t)j * W + c];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
        if (mx - mn > 1e-12 * (fabs(mx) + fabs(mn) + 1e-300)) use.push_back((int)c);
    }
    const size_t k = use.size();
    // centre and scale each used feature for conditioning
    std::vector<double> mu(k, 0.0), sd(k, 1.0);
    for (size_t a = 1; a < k; ++a) {
        double s = 0, s2 = 0;
        for (uint32_t j = 0; j < m; ++j) s += rows[(size_t)j * W + use[a]];
        mu[a] = s / m;
        for (uint32_t j = 0; j < m; ++j) {
            const double v = rows[(size_t)j * W + use[a]] - mu[a];
            s2 += v * v;
        }
        sd[a] = sqrt(s2 / m);
        if (!(sd[a] > 0)) sd[a] = 1.0;
    }
    std::vector<double> A(k * k, 0.0), b(k, 0.0), x(k);
    for (uint32_t j = 0; j < m; ++j) {
        const double *r = &rows[(size_t)j * W];
        for (size_t a = 0; a < k; ++a) x[a] = a == 0 ? 1.0 : (r[use[a]] - mu[a]) / sd[a];
        for (size_t a = 0; a < k; ++a) {
            b[a] += x[a] * r[P + 2];
            for (size_t c = 0; c < k; ++c) A[a * k + c] += x[a] * x[c];
        }
    }
    for (size_t a = 1; a < k; ++a) A[a * k + a] += 1e-9 * m;
    // Cholesky solve
    std::vector<double> L(k * k, 0.0);
    for (size_t i = 0; i < k; ++i)
        for (size_t j = 0; j <= i; ++j) {
            double s = A[i * k + j];
            for (size_t q = 0; q < j; ++q) s -= L[i * k + q] * L[j * k + q];
            L[i * k + j] = i == j ? sqrt(s > 1e-300 ? s : 1e-300) : s / L[j * k + j];
        }
    std::vector<double> z(k), beta(k);
    for (size_t i = 0; i < k; ++i) {
        double s = b[i];
        for (size_t q = 0; q < i; ++q) s -= L[i * k + q] * z[q];
        z[i] = s / L[i * k + i];
    }
    for (size_t i = k; i-- > 0;) {
        double s = z[i];
        for (size_t q = i + 1; q < k; ++q) s -= L[q * k + i] * beta[q];
        beta[i] = s / L[i * k + i];
    }
    // back to raw feature units: coef[1 + c] for feature column c
    memset(tn.coef, 0, sizeof tn.coef);
    double c0 = beta[0];
    for (size_t a = 1; a < k; ++a) {
        const double w = beta[a] / sd[a];
        tn.coef[1 + use[a]] = w;
        c0 -= w * mu[a];
    }
    tn.coef[0] = c0;
}

static int timed_solve(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                       const bg_launch &lc, hipStream_t stream, double &ms) {
    hipEvent_t e0, e1;
    BG_HIP_CHECK(hipEventCreate(&e0));
    BG_HIP_CHECK(hipEventCreate(&e1));
    BG_HIP_CHECK(hipEventRecord(e0, stream));
    int rc = rk_launch_ensemble(method, rhs, semantics, d, lc, stream);
    if (rc == BG_OK) {
        BG_HIP_CHECK(hipEventRecord(e1, stream));
        BG_HIP_CHECK(hipEventSynchronize(e1));
        float f = 0;
        BG_HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
        ms = f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int rk_autotune(int method, int rhs, int semantics, const bg_ensemble_desc &din,
                bg_tuning &tn, hipStream_t stream) {
    memset(&tn, 0, sizeof tn);
    if (!rk_supported(method, rhs, din.dim) && rhs != BG_RHS_LINEAR) return BG_ERR_UNSUPPORTED;
    const uint64_t n = din.n;
    const uint32_t P = din.n_params < BG_TUNE_MAX_PARAMS ? din.n_params : BG_TUNE_MAX_PARAMS;
    tn.n_params = P;
    if (n < 2 || rhs == BG_RHS_LINEAR) return BG_OK;  // defaults: nothing to schedule
    if (n > 0x7fffffffull || !din.y0 || (P && !din.params)) return BG_ERR_INVALID_ARGUMENT;
    if (din.order) return BG_ERR_INVALID_ARGUMENT;   // the tuner owns the schedule
    // inside the do/while below: record the failure and fall through to the
    // frees after the loop
#define BG_TRY(expr)                  \
    if ((expr) != hipSuccess) {       \
        rc = BG_ERR_HIP;              \
        break;                        \
    }

    bg_ensemble_desc d = din;
    // step counts: the caller's buffer or a scratch one
    bg_counts *scratch_counts = nullptr;
    if (!d.counts) {
        BG_HIP_CHECK(hipMallocAsync((void **)&scratch_counts, n * sizeof(bg_counts), stream));
        d.counts = scratch_counts;
    }
    int32_t *order = nullptr;
    double *rows_d = nullptr;
    int rc = BG_OK;
    do {
        // 1. a solve in input order with the default shape gives each
        //    instance's attempt count
        bg_launch lc0;
        memset(&lc0, 0, sizeof lc0);
        double ms0 = 0;
        if ((rc = timed_solve(method, rhs, semantics, d, lc0, stream, ms0)) != BG_OK) break;
        tn.pilot_ms = ms0;
        // 2. fit the cost model on a deterministic subsample
        const uint32_t m = (uint32_t)(n < 16384 ? n : 16384);
        const uint64_t stride = n / m;
        const uint32_t W = P + 3;
        BG_TRY(hipMallocAsync((void **)&rows_d, (size_t)m * W * sizeof(double), stream));
        tune_sample_kernel<<<(m + 255) / 256, 256, 0, stream>>>(d, d.counts, stride, m, P, rows_d);
        BG_TRY(hipGetLastError());
        std::vector<double> rows((size_t)m * W);
        BG_TRY(hipMemcpyAsync(rows.data(), rows_d, rows.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, stream));
        BG_TRY(hipStreamSynchronize(stream));
        tn.pilot_n = m;
        fit_model(rows, m, P, tn);
        // 3. schedule, then time the candidate shapes on the whole problem
        BG_TRY(hipMallocAsync((void **)&order, n * sizeof(int32_t), stream));
        if ((rc = schedule(tn, d, order, stream)) != BG_OK) break;
        d.order = order;
        const uint32_t cand[][2] = {{256, 0}, {256, 4}, {256, 3}, {256, 2}, {128, 6}, {64, 12}};
        double best = 1e300;
        for (const auto &c : cand) {
            bg_launch lc;
            memset(&lc, 0, sizeof lc);
            lc.block = c[0];
            lc.blocks_per_cu = c[1];
            double ms = 0;
            if ((rc = timed_solve(method, rhs, semantics, d, lc, stream, ms)) != BG_OK) break;
            ++tn.n_trials;
            if (c[1] == 0 && c[0] == 256) tn.default_ms = ms;
            if (ms < best) {
                best = ms;
                tn.launch = lc;
            }
        }
        if (rc != BG_OK) break;
        tn.best_ms = best;
    } while (0);
#undef BG_TRY
    if (rows_d) hipFreeAsync(rows_d, stream);
    if (order) hipFreeAsync(order, stream);
    if (scratch_counts) hipFreeAsync(scratch_counts, stream);
    const hipError_t e = hipStreamSynchronize(stream);
    if (rc == BG_OK && e != hipSuccess) rc = BG_ERR_HIP;
    return rc;
}

int rk_schedule(const bg_tuning &tn, const bg_ensemble_desc &d, int32_t *order,
                hipStream_t stream) {
    return schedule(tn, d, order, stream);
}

// Keys of the measured schedule: an earlier solve's attempts per instance
// (accepted + rejected, saturating), instance indices beside them.
__global__ void count_keys_kernel(const bg_counts *__restrict__ counts, uint64_t n,
                                  uint32_t *__restrict__ keys, int32_t *__restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bg_counts c = counts[i];
    const uint32_t a = c.n_accepted + c.n_rejected;
    keys[i] = a < c.n_accepted ? 0xffffffffu : a;
    idx[i] = (int32_t)i;
}

// Longest first by measured work: a stable descending radix sort of the
// keys, so equal counts keep index order and the permutation is
// deterministic.
int schedule_by_counts(const bg_counts *counts, uint64_t n, int32_t *order, hipStream_t stream) {
    if (n == 0) return BG_OK;
    if (n > 0x7fffffffull || !counts || !order) return BG_ERR_INVALID_ARGUMENT;
    size_t sort_bytes = 0;
    BG_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, (const uint32_t *)nullptr,
                                                (uint32_t *)nullptr, (const int32_t *)nullptr,
                                                (int32_t *)nullptr, (uint32_t)n, 0, 32, stream));
    const size_t kbytes = ((n * sizeof(uint32_t)) + 255) & ~(size_t)255;
    char *ws = nullptr;
    BG_HIP_CHECK(hipMallocAsync((void **)&ws, 3 * kbytes + sort_bytes, stream));
    uint32_t *keys = (uint32_t *)ws;
    uint32_t *keys_sorted = (uint32_t *)(ws + kbytes);
    int32_t *idx = (int32_t *)(ws + 2 * kbytes);
    void *tmp = ws + 3 * kbytes;
    count_keys_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, stream>>>(counts, n, keys, idx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs_desc(tmp, sort_bytes, keys, keys_sorted, idx, order,
                                           (uint32_t)n, 0, 32, stream);
    // the workspace goes back on every path
    BG_HIP_CHECK(hipFreeAsync(ws, stream));
    BG_HIP_CHECK(e);
    return BG_OK;
}

static int timed_bdf(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                     uint32_t cap, hipStream_t stream, double &ms) {
    hipEvent_t e0, e1;
    BG_HIP_CHECK(hipEventCreate(&e0));
    BG_HIP_CHECK(hipEventCreate(&e1));
    BG_HIP_CHECK(hipEventRecord(e0, stream));
    int rc = bdf_launch_ensemble(method, rhs, semantics, lu_fp32, d, stream, nullptr, 0, cap);
    if (rc == BG_OK) {
        BG_HIP_CHECK(hipEventRecord(e1, stream));
        BG_HIP_CHECK(hipEventSynchronize(e1));
        float f = 0;
        BG_HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
        ms = f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int bdf_autotune(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &din,
                 bg_tuning &tn, hipStream_t stream) {
    memset(&tn, 0, sizeof tn);
    tn.launch.block = 64;
    if (!bdf_supported(method, rhs, din.dim)) return BG_ERR_UNSUPPORTED;
    const uint64_t n = din.n;
    const uint32_t P = din.n_params < BG_TUNE_MAX_PARAMS ? din.n_params : BG_TUNE_MAX_PARAMS;
    tn.n_params = P;
    if (n < 2) return BG_OK;
    if (n > 0x7fffffffull || !din.y0 || (P && !din.params)) return BG_ERR_INVALID_ARGUMENT;
    if (din.order) return BG_ERR_INVALID_ARGUMENT;  // the tuner owns the schedule
#define BG_TRY(expr)                  \
    if ((expr) != hipSuccess) {       \
        rc = BG_ERR_HIP;              \
        break;                        \
    }
    bg_ensemble_desc d = din;
    bg_counts *scratch_work = nullptr;
    if (!d.work) {
        BG_HIP_CHECK(hipMallocAsync((void **)&scratch_work, n * sizeof(bg_counts), stream));
        d.work = scratch_work;
    }
    int32_t *order = nullptr;
    double *rows_d = nullptr;
    int rc = BG_OK;
    do {
        // 1. a solve in input order at the occupancy limit gives the work
        double ms0 = 0;
        if ((rc = timed_bdf(method, rhs, semantics, lu_fp32, d, 0, stream, ms0)) != BG_OK) break;
        tn.pilot_ms = ms0;
        // 2. the cost model on a deterministic subsample
        const uint32_t m = (uint32_t)(n < 16384 ? n : 16384);
        const uint64_t stride = n / m;
        const uint32_t W = P + 3;
        BG_TRY(hipMallocAsync((void **)&rows_d, (size_t)m * W * sizeof(double), stream));
        tune_sample_bdf_kernel<<<(m + 255) / 256, 256, 0, stream>>>(d, d.work, stride, m, P, rows_d);
        BG_TRY(hipGetLastError());
        std::vector<double> rows((size_t)m * W);
        BG_TRY(hipMemcpyAsync(rows.data(), rows_d, rows.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        BG_TRY(hipStreamSynchronize(stream));
        tn.pilot_n = m;
        fit_model(rows, m, P, tn);
        // 3. schedule, then the resident-wave caps
        BG_TRY(hipMallocAsync((void **)&order, n * sizeof(int32_t), stream));
        if ((rc = schedule(tn, d, order, stream)) != BG_OK) break;
        d.order = order;
        // caps of 4, 3, 2 and 1 waves per SIMD below the occupancy limit
        size_t ws_bytes = 0;
        int occ = 0;
        if ((rc = bdf_workspace_bytes(method, rhs, semantics, lu_fp32, d, &ws_bytes, &occ)) !=
            BG_OK)
            break;
        double best = 1e300;
        for (const uint32_t cap : {0u, 16u, 12u, 8u, 4u}) {
            if (cap && (int)cap >= occ) continue;
            double ms = 0;
            if ((rc = timed_bdf(method, rhs, semantics, lu_fp32, d, cap, stream, ms)) != BG_OK)
                break;
            ++tn.n_trials;
            if (cap == 0) tn.default_ms = ms;
            if (ms < best) {
                best = ms;
                tn.launch.blocks_per_cu = cap;
            }
        }
        if (rc != BG_OK) break;
        tn.best_ms = best;
    } while (0);
#undef BG_TRY
    if (rows_d) hipFreeAsync(rows_d, stream);
    if (order) hipFreeAsync(order, stream);
    if (scratch_work) hipFreeAsync(scratch_work, stream);
    const hipError_t e = hipStreamSynchronize(stream);
    if (rc == BG_OK && e != hipSuccess) rc = BG_ERR_HIP;
    return rc;
}

int bdf_solve_tuned(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &din,
                    const bg_tuning &tn, hipStream_t stream) {
    const uint32_t cap = tn.launch.blocks_per_cu;
    if (din.order || din.n < 2)
        return bdf_launch_ensemble(method, rhs, semantics, lu_fp32, din, stream, nullptr, 0, cap);
    bg_ensemble_desc d = din;
    int32_t *order = nullptr;
    BG_HIP_CHECK(hipMallocAsync((void **)&order, d.n * sizeof(int32_t), stream));
    int rc = schedule(tn, d, order, stream);
    if (rc == BG_OK) {
        d.order = order;
        rc = bdf_launch_ensemble(method, rhs, semantics, lu_fp32, d, stream, nullptr, 0, cap);
    }
    BG_HIP_CHECK(hipFreeAsync(order, stream));
    return rc;
}

// The plan's gradient solve: its K, capacity and blocks per CU, and the
// fitted forward order when it has one and the caller gave none.
static int adjoint_run_plan(int method, int rhs, const bg_adjoint_desc &din, uint32_t K,
                            uint32_t maxck, uint32_t bpc, const bg_tuning *model,
                            hipStream_t stream) {
    bg_adjoint_desc d = din;
    d.checkpoint_every = K;
    d.max_checkpoints = maxck;
    d.blocks_per_cu = bpc;
    if (!model || d.base.order || d.base.n < 2) return rk_launch_adjoint(method, rhs, d, stream);
    int32_t *order = nullptr;
    BG_HIP_CHECK(hipMallocAsync((void **)&order, d.base.n * sizeof(int32_t), stream));
    int rc = schedule(*model, d.base, order, stream);
    if (rc == BG_OK) {
        d.base.order = order;
        rc = rk_launch_adjoint(method, rhs, d, stream);
    }
    BG_HIP_CHECK(hipFreeAsync(order, stream));
    return rc;
}

int adjoint_solve_tuned(int method, int rhs, const bg_adjoint_desc &d,
                        const bg_adjoint_tuning &plan, hipStream_t stream) {
    if (plan.schedule && plan.forward.n_params > BG_TUNE_MAX_PARAMS)
        return BG_ERR_INVALID_ARGUMENT;
    return adjoint_run_plan(method, rhs, d, plan.checkpoint_every, plan.max_checkpoints,
                            plan.blocks_per_cu, plan.schedule ? &plan.forward : nullptr, stream);
}

int adjoint_autotune(int method, int rhs, const bg_adjoint_desc &din, bg_adjoint_tuning &tn,
                     hipStream_t stream) {
    memset(&tn, 0, sizeof tn);
    const uint64_t n = din.base.n;
    const uint32_t K0 = din.checkpoint_every ? din.checkpoint_every : 16u;
    const uint32_t M0 = din.max_checkpoints ? din.max_checkpoints : 128u;
    const uint64_t steps = (uint64_t)K0 * M0;  // accepted steps the caller's plan covers
    struct Trial {
        uint32_t K, bpc;
    };
    Trial trials[12];
    int nt = 0;
    for (const uint32_t K : {K0, 16u, 32u, 48u})
        for (const uint32_t bpc : {din.blocks_per_cu, 0u, 1u, 2u}) {
            bool seen = false;
            for (int q = 0; q < nt; ++q) seen |= trials[q].K == K && trials[q].bpc == bpc;
            if (!seen && nt < 12) trials[nt++] = Trial{K, bpc};
        }
    // the forward order is tuned too unless the caller owns it
    const uint32_t P = din.base.n_params < BG_TUNE_MAX_PARAMS ? din.base.n_params
                                                              : BG_TUNE_MAX_PARAMS;
    const bool fit = !din.base.order && n >= 2 && (P == 0 || din.base.params);
    bg_adjoint_desc dp = din;  // the pilot: the caller's counts or a scratch buffer
    bg_counts *scratch_counts = nullptr;
    if (fit && !dp.base.counts) {
        BG_HIP_CHECK(hipMallocAsync((void **)&scratch_counts, n * sizeof(bg_counts), stream));
        dp.base.counts = scratch_counts;
    }
    double *rows_d = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BG_OK;
    double best = 1e300;
    do {
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
        // every (K, blocks per CU) in the caller's order, then (when fitted)
        // in the fitted order; trial 0 is the caller's own plan, whose
        // attempt counts fit the model
        for (int pass = 0; pass < (fit ? 2 : 1) && rc == BG_OK; ++pass) {
            if (pass == 1) {
                const uint32_t m = (uint32_t)(n < 16384 ? n : 16384);
                const uint64_t stride = n / m;
                const uint32_t W = P + 3;
                if (hipMallocAsync((void **)&rows_d, (size_t)m * W * sizeof(double), stream) !=
                    hipSuccess) {
                    rc = BG_ERR_HIP;
                    break;
                }
                tune_sample_kernel<<<(m + 255) / 256, 256, 0, stream>>>(dp.base, dp.base.counts,
                                                                       stride, m, P, rows_d);
                std::vector<double> rows((size_t)m * W);
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(rows.data(), rows_d, rows.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, stream) != hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess) {
                    rc = BG_ERR_HIP;
                    break;
                }
                tn.forward.n_params = P;
                tn.forward.pilot_n = m;
                fit_model(rows, m, P, tn.forward);
            }
            for (int q = 0; q < nt; ++q) {
                const uint32_t K = trials[q].K;
                const uint32_t maxck = (uint32_t)((steps + K - 1) / K);
                if (hipEventRecord(e0, stream) != hipSuccess) {
                    rc = BG_ERR_HIP;
                    break;
                }
                rc = adjoint_run_plan(method, rhs, pass == 0 && q == 0 ? dp : din, K, maxck,
                                      trials[q].bpc, pass == 1 ? &tn.forward : nullptr, stream);
                if (rc != BG_OK) break;
                float ms = 0;
                if (hipEventRecord(e1, stream) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = BG_ERR_HIP;
                    break;
                }
                ++tn.n_trials;
                if (pass == 0 && q == 0) tn.default_ms = ms;  // the caller's own plan
                if (ms < best) {
                    best = ms;
                    tn.checkpoint_every = K;
                    tn.max_checkpoints = maxck;
                    tn.blocks_per_cu = trials[q].bpc;
                    tn.schedule = pass == 1 ? 1u : 0u;
                }
            }
        }
    } while (0);
    tn.best_ms = tn.n_trials ? best : 0.0;
    if (rows_d) hipFreeAsync(rows_d, stream);
    if (scratch_counts) hipFreeAsync(scratch_counts, stream);
    const hipError_t e = hipStreamSynchronize(stream);
    if (rc == BG_OK && e != hipSuccess) rc = BG_ERR_HIP;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return rc;
}

int rk_solve_tuned(int method, int rhs, int semantics, const bg_ensemble_desc &din,
                   const bg_tuning &tn, hipStream_t stream) {
    if (rhs == BG_RHS_LINEAR || din.order || din.n < 2)
        return rk_launch_ensemble(method, rhs, semantics, din, tn.launch, stream);
    bg_ensemble_desc d = din;
    int32_t *order = nullptr;
    BG_HIP_CHECK(hipMallocAsync((void **)&order, d.n * sizeof(int32_t), stream));
    int rc = schedule(tn, d, order, stream);
    if (rc == BG_OK) {
        d.order = order;
        rc = rk_launch_ensemble(method, rhs, semantics, d, tn.launch, stream);
    }
    BG_HIP_CHECK(hipFreeAsync(order, stream));
    return rc;
}

}  // namespace bg
