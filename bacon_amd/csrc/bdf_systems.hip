// bdf_systems.hip — batched BDF (BDF6 / BDF2) for small systems of any
// dimension up to 16 (bacon-sci's BDF is generic over the dimension,
// src/ivp/bdf.rs:144-170, 334-634): Lotka-Volterra and the harmonic
// oscillator (2), Lorenz-63 and Robertson (3) as gathered register
// right-hand sides in groups of 2 and 4 lanes, and the Brusselator with 2,
// 4 and 8 cells (4, 8, 16 components) in groups of D lanes.  Reached through
// the entry points of bdf_kernels.hip; the device code is bdf_device.h.
#include "bdf_device.h"

namespace bg {

template <class V>
static int visit_sys(int rhs, uint32_t dim, V &&v) {
    switch (rhs) {
    case BG_RHS_LOTKA_VOLTERRA:
        return dim == 2 ? v(BTag<GatherBdf<RhsLotkaVolterra>>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_OSCILLATOR:
        return dim == 2 ? v(BTag<GatherBdf<RhsOscillator>>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_LORENZ:
        return dim == 3 ? v(BTag<GatherBdf<RhsLorenz>>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_ROBERTSON:
        return dim == 3 ? v(BTag<GatherBdf<RhsRobertson>>{}) : BG_ERR_UNSUPPORTED;
    case BG_RHS_BRUSSELATOR:
        if (dim == 4) return v(BTag<RhsBrusselatorG<2>>{});
        if (dim == 8) return v(BTag<RhsBrusselatorG<4>>{});
        if (dim == 16) return v(BTag<RhsBrusselatorG<8>>{});
        return BG_ERR_UNSUPPORTED;
    default:
        return BG_ERR_UNSUPPORTED;
    }
}

// The Jacobian inverse of these systems is FP64 (lu_fp32 is accepted and
// ignored, as for the 1-D systems).
template <class V>
static int visit_bdf(int method, int rhs, int semantics, uint32_t dim, V &&v) {
    return visit_sys(rhs, dim, [&](auto rtag) -> int {
        return bdf_with_order_sem(method, semantics,
                                  [&](auto otag, auto stag) -> int { return v(rtag, otag, stag); });
    });
}

bool bdf_sys_supported(int method, int rhs, uint32_t dim) {
    return visit_bdf(method, rhs, BG_INTENDED, dim, [](auto, auto, auto) { return BG_OK; }) == BG_OK;
}

int bdf_sys_workspace_bytes(int method, int rhs, int semantics, int, const bg_ensemble_desc &d,
                            size_t *bytes, int *occupancy) {
    return visit_bdf(method, rhs, semantics, d.dim, [&](auto rtag, auto otag, auto stag) -> int {
        using R = typename decltype(rtag)::type;
        using S = typename decltype(stag)::type;
        constexpr int O = decltype(otag)::value;
        uint64_t grid = 0;
        return bdf_launch_shape<R, O, S, double>(d, grid, *bytes, 0, occupancy);
    });
}

int bdf_sys_launch_ensemble(int method, int rhs, int semantics, int, const bg_ensemble_desc &d,
                            hipStream_t stream, void *ws_in, size_t ws_bytes, uint32_t cap_per_cu) {
    return visit_bdf(method, rhs, semantics, d.dim, [&](auto rtag, auto otag, auto stag) -> int {
        using R = typename decltype(rtag)::type;
        using S = typename decltype(stag)::type;
        constexpr int O = decltype(otag)::value;
        return bdf_launch_one<R, O, S, double>(d, stream, ws_in, ws_bytes, cap_per_cu);
    });
}

int bdf_sys_launch_path(int method, int rhs, int semantics, uint32_t dim, BdfPathState *state,
                        double *path_t, double *path_y, uint32_t cap, hipStream_t stream) {
    return visit_bdf(method, rhs, semantics, dim, [&](auto rtag, auto otag, auto stag) -> int {
        using R = typename decltype(rtag)::type;
        using S = typename decltype(stag)::type;
        constexpr int O = decltype(otag)::value;
        return bdf_path_one<R, O, S, double>(state, path_t, path_y, cap, stream);
    });
}

}  // namespace bg
