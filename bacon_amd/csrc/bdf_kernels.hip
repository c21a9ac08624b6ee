// bdf_kernels.hip — batched BDF (BDF6 / BDF2) for gfx950: the 1-D reference
// test right-hand sides and the 32-component Brusselator (one instance per
// wave, FP64 or FP32 Jacobian inverse), and the entry points of the BDF
// family (the small systems of bdf_systems.hip are reached through them).
//
// Hot path: bacon-sci's BDFSolver::step (src/ivp/bdf.rs:495-634) with its
// RK4 bootstrap (bdf.rs:346-387), finite-difference Jacobian (bdf.rs:390-411)
// and Broyden corrector (bdf.rs:414-475); the device code is bdf_device.h.
#include "bdf_device.h"

namespace bg {

// system -> (right-hand-side adaptor, inverse arithmetic)
template <class V>
static int visit_sys(int rhs, uint32_t dim, int lu_fp32, V &&v) {
    switch (rhs) {
    case BG_RHS_EXP:
        return dim == 1 ? v(BTag<ScalarBdf<RhsExp>>{}, BTag<double>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_NEG:
        return dim == 1 ? v(BTag<ScalarBdf<RhsNeg>>{}, BTag<double>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_QUADRATIC:
        return dim == 1 ? v(BTag<ScalarBdf<RhsQuadratic>>{}, BTag<double>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_COS:
        return dim == 1 ? v(BTag<ScalarBdf<RhsCos>>{}, BTag<double>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_FAIL_AFTER:
        return dim == 1 ? v(BTag<ScalarBdf<RhsFailAfter>>{}, BTag<double>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_BRUSSELATOR: {
        if (dim != 32) return BG_ERR_UNSUPPORTED;
        auto rt = BTag<RhsBrusselatorG<16>>{};
        return lu_fp32 ? v(rt, BTag<float>{}) : v(rt, BTag<double>{});
    }
    default:
        return BG_ERR_UNSUPPORTED;
    }
}

template <class V>
static int visit_bdf(int method, int rhs, int semantics, uint32_t dim, int lu_fp32, V &&v) {
    return visit_sys(rhs, dim, lu_fp32, [&](auto rtag, auto ftag) -> int {
        return bdf_with_order_sem(method, semantics,
                                  [&](auto otag, auto stag) -> int { return v(rtag, otag, stag, ftag); });
    });
}

bool bdf_supported(int method, int rhs, uint32_t dim) {
    return visit_bdf(method, rhs, BG_INTENDED, dim, 0,
                     [](auto, auto, auto, auto) { return BG_OK; }) == BG_OK ||
           bdf_sys_supported(method, rhs, dim);
}

int bdf_workspace_bytes(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                        size_t *bytes, int *occupancy) {
    const int rc = visit_bdf(method, rhs, semantics, d.dim, lu_fp32,
                             [&](auto rtag, auto otag, auto stag, auto ftag) -> int {
                                 using R = typename decltype(rtag)::type;
                                 using S = typename decltype(stag)::type;
                                 using F = typename decltype(ftag)::type;
                                 constexpr int O = decltype(otag)::value;
                                 uint64_t grid = 0;
                                 return bdf_launch_shape<R, O, S, F>(d, grid, *bytes, 0,
                                                                     occupancy);
                             });
    if (rc != BG_ERR_UNSUPPORTED) return rc;
    return bdf_sys_workspace_bytes(method, rhs, semantics, lu_fp32, d, bytes, occupancy);
}

int bdf_launch_ensemble(int method, int rhs, int semantics, int lu_fp32, const bg_ensemble_desc &d,
                        hipStream_t stream, void *ws_in, size_t ws_bytes, uint32_t cap_per_cu) {
    const int rc = visit_bdf(method, rhs, semantics, d.dim, lu_fp32,
                             [&](auto rtag, auto otag, auto stag, auto ftag) -> int {
                                 using R = typename decltype(rtag)::type;
                                 using S = typename decltype(stag)::type;
                                 using F = typename decltype(ftag)::type;
                                 constexpr int O = decltype(otag)::value;
                                 return bdf_launch_one<R, O, S, F>(d, stream, ws_in, ws_bytes,
                                                                   cap_per_cu);
                             });
    if (rc != BG_ERR_UNSUPPORTED) return rc;
    return bdf_sys_launch_ensemble(method, rhs, semantics, lu_fp32, d, stream, ws_in, ws_bytes,
                                   cap_per_cu);
}

int bdf_launch_inverse(uint32_t dim, uint64_t count, const double *a, double *inv,
                       int32_t *branch, int first, hipStream_t stream) {
    if (first < 0 || first > 2) return BG_ERR_INVALID_ARGUMENT;
    if (count == 0) return BG_OK;
    if (!a || !inv || !branch || count > 0x7fffffffull) return BG_ERR_INVALID_ARGUMENT;
    auto run = [&](auto kern, int per_block) -> int {
        const uint64_t blocks = (count + per_block - 1) / per_block;
        kern<<<(uint32_t)blocks, 64, 0, stream>>>(a, inv, branch, count, first);
        BG_HIP_CHECK(hipGetLastError());
        return BG_OK;
    };
    switch (dim) {
    case 1: return run(bdf_inverse_kernel<1, 1>, 64);
    case 2: return run(bdf_inverse_kernel<2, 1>, 32);
    case 4: return run(bdf_inverse_kernel<4, 1>, 16);
    case 8: return run(bdf_inverse_kernel<8, 1>, 8);
    case 16: return run(bdf_inverse_kernel<16, 1>, 4);
    case 32: return run(bdf_inverse_kernel<32, 2>, 1);
    default: return BG_ERR_UNSUPPORTED;
    }
}

int bdf_launch_path(int method, int rhs, int semantics, uint32_t dim, BdfPathState *state,
                    double *path_t, double *path_y, uint32_t cap, hipStream_t stream) {
    const int rc = visit_bdf(method, rhs, semantics, dim, 0,
                             [&](auto rtag, auto otag, auto stag, auto ftag) -> int {
                                 using R = typename decltype(rtag)::type;
                                 using S = typename decltype(stag)::type;
                                 using F = typename decltype(ftag)::type;
                                 constexpr int O = decltype(otag)::value;
                                 return bdf_path_one<R, O, S, F>(state, path_t, path_y, cap, stream);
                             });
    if (rc != BG_ERR_UNSUPPORTED) return rc;
    return bdf_sys_launch_path(method, rhs, semantics, dim, state, path_t, path_y, cap, stream);
}

}  // namespace bg
