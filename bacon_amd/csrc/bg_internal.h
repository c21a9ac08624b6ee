// bg_internal.h — host-side interfaces shared by the translation units of
// libbacon_gpu.so (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bacon_gpu.h"

namespace bg {

// Device workspace of one persistent launch: the work counter at [0, 256),
// the descriptor copy after it.  A launcher given ws == nullptr takes it from
// the stream-ordered allocator and frees it on the stream; a caller-owned
// workspace (graph capture) must hold kLaunchWorkspaceBytes.
constexpr size_t kLaunchWorkspaceBytes = 256 + ((sizeof(bg_ensemble_desc) + 255) & ~(size_t)255);

// Launches of the RK family.  Each returns a BG_* status.
int rk_launch_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                       const bg_launch &launch, hipStream_t stream, void *ws = nullptr);
int rk_launch_plain(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                    hipStream_t stream);
bool rk_supported(int method, int rhs, uint32_t dim);
int rk_launch_path_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                            const bg_path_desc &p, hipStream_t stream);

// Adams family (adams_kernels.hip)
bool adams_supported(int method, int rhs, uint32_t dim);
int adams_launch_ensemble(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                          hipStream_t stream, void *ws = nullptr);
// Single-instance Adams path state (resumable, like PathState below); the
// Adams right-hand sides have at most 3 components.
struct AdamsPathState {
    double t, dt, end, tol, dt_min, dt_max;
    uint64_t steps, max_steps;
    uint32_t n_accepted, n_rejected;
    int32_t vlen, ym, status, finished;
    uint32_t count, pad;
    double y[4], save[4], impl[4];
    double dq[4][4];
    double vt[4], vy[4][4];
    double p[8];
};
int adams_launch_path(int method, int rhs, int semantics, uint32_t dim, AdamsPathState *state,
                      double *path_t, double *path_y, uint32_t cap, hipStream_t stream);

// Autotuner (tune.hip)
int rk_autotune(int method, int rhs, int semantics, const bg_ensemble_desc &d, bg_tuning &tn,
                hipStream_t stream);
int rk_schedule(const bg_tuning &tn, const bg_ensemble_desc &d, int32_t *order,
                hipStream_t stream);
int rk_solve_tuned(int method, int rhs, int semantics, const bg_ensemble_desc &d,
                   const bg_tuning &tn, hipStream_t stream);
int schedule_by_counts(const bg_counts *counts, uint64_t n, int32_t *order, hipStream_t stream);
// BDF: the same fitted schedule on the work counters, and the resident
// waves per CU (tn.launch.blocks_per_cu; blocks are single waves)
int bdf_autotune(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                 bg_tuning &tn, hipStream_t stream);
int bdf_solve_tuned(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                    const bg_tuning &tn, hipStream_t stream);

// RK45 for y' = A y + b with one shared A (dim 32, 64 or 128), the stage
// right-hand sides as FP64 MFMA GEMMs over tiles of 16 instances
// (linear_kernels.hip).
bool linear_supported(uint32_t dim);
int linear_launch_ensemble(int method, int semantics, const bg_ensemble_desc &d,
                           hipStream_t stream, void *ws = nullptr);

// Single-instance path production (one lane).  The state block lives in
// device memory and is resumable: the kernel stops when the path buffer is
// full and the host relaunches it after draining the buffer.
struct PathState {
    double t, dt, end, tol, dt_min, dt_max;
    uint64_t attempts, max_attempts;
    uint32_t n_accepted, n_rejected;
    int32_t status;   // BG_OK while running / at Done, else the failure code
    int32_t finished; // 1 once Done or failed
    uint32_t count;   // entries written into the path buffer by this launch
    uint32_t pad;
    double y[128];
    double K[6][128];
    double p[8];
};
int rk_launch_path(int method, int rhs, int semantics, uint32_t dim, PathState *state,
                   double *path_t, double *path_y, uint32_t cap, hipStream_t stream);

// BDF family (bdf_kernels.hip)
struct BdfPathState {
    double t, dt, end, tol, dt_min, dt_max;
    uint64_t steps, max_steps;
    uint32_t n_accepted, n_rejected, n_newton, n_lu;
    int32_t yield_memory, hist_full, status, finished;
    uint32_t count, pad;
    double y[64], save[64];
    double hist[7][64];
    double hist_t[7];
    double p[8];
};
bool bdf_supported(int method, int rhs, uint32_t dim);
// ws == nullptr: the launcher allocates the workspace; else ws holds
// ws_bytes, at least bdf_workspace_bytes() for the same arguments, which
// also gives the kernel's resident waves (blocks) per CU in *occupancy.
// cap_per_cu > 0 caps the resident waves per CU.
int bdf_launch_ensemble(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                        hipStream_t stream, void *ws = nullptr, size_t ws_bytes = 0,
                        uint32_t cap_per_cu = 0);
int bdf_workspace_bytes(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                        size_t *bytes, int *occupancy = nullptr);
int bdf_launch_path(int method, int rhs, int semantics, uint32_t dim, BdfPathState *state,
                    double *path_t, double *path_y, uint32_t cap, hipStream_t stream);
// the corrector's Jacobian inverse cascade on caller matrices
int bdf_launch_inverse(uint32_t dim, uint64_t count, const double *a, double *inv,
                       int32_t *branch, int first, hipStream_t stream);
// the small systems (bdf_systems.hip), reached through the functions above
bool bdf_sys_supported(int method, int rhs, uint32_t dim);
int bdf_sys_workspace_bytes(int method, int rhs, int semantics, int lu_fp32,
                            const bg_ensemble_desc &d, size_t *bytes, int *occupancy);
int bdf_sys_launch_ensemble(int method, int rhs, int semantics, int lu_fp32,
                            const bg_ensemble_desc &d, hipStream_t stream, void *ws_in,
                            size_t ws_bytes, uint32_t cap_per_cu);
int bdf_sys_launch_path(int method, int rhs, int semantics, uint32_t dim, BdfPathState *state,
                        double *path_t, double *path_y, uint32_t cap, hipStream_t stream);

// adjoint (adjoint_kernels.hip), and its tuner (tune.hip)
int rk_launch_adjoint(int method, int rhs, const bg_adjoint_desc &d, hipStream_t stream);
int adjoint_solve_tuned(int method, int rhs, const bg_adjoint_desc &d,
                        const bg_adjoint_tuning &plan, hipStream_t stream);
int adjoint_autotune(int method, int rhs, const bg_adjoint_desc &d, bg_adjoint_tuning &out,
                     hipStream_t stream);

int device_cu_count(int device);

}  // namespace bg

#define BG_HIP_CHECK(expr)                         \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return BG_ERR_HIP;   \
    } while (0)
