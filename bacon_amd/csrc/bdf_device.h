// bdf_device.h — device side of the batched BDF (BDF6 / BDF2) for gfx950,
// shared by bdf_kernels.hip and bdf_systems.hip.
//
// Hot path: bacon-sci's BDFSolver::step (src/ivp/bdf.rs:495-634) with its
// RK4 bootstrap (bdf.rs:346-387), finite-difference Jacobian (bdf.rs:390-411)
// and Broyden corrector (bdf.rs:414-475), recast as an ensemble.
//
// Layout: one instance per group of W lanes, the power of two W >= D (the
// whole wave for the 32-component Brusselator, split into half rows; one
// lane for scalar problems; W - D padding lanes for small systems such as
// Lorenz-63).  Lane g of a group
// owns component g of every vector (state, guess, residuals, the history of
// the last O accepted values) and row g of the Jacobian inverse, all in VGPRs.
// Group collectives use cross-lane shuffles (ds_bpermute); the column sums of
// the Broyden update u = s^T J^-1 go through an LDS transpose (row stride
// W+1 doubles: conflict-free column reads).  The inverse is Gauss-Jordan with
// partial pivoting on the rows (implicit, no row swaps); its arithmetic type
// is a template parameter (double, or float with the Broyden iteration and
// all residuals in double).  Instances are refilled per group from a global
// work counter, like the RK ensemble kernel.
#pragma once

#include <type_traits>

#include "bg_device.h"
#include "bg_internal.h"

namespace bg {

constexpr uint32_t kBdfDefaultMaxSteps = 1000000u;


// ---------------------------------------------------------------------------
// cross-lane moves without LDS traffic (gfx950 lane layouts, measured):
//   DPP row_ror:1 (0x121): lane i <- lane (i - 1) mod 16 within its row of 16
//   DPP row_ror:15 (0x12F): lane i <- lane (i + 1) mod 16
//   DPP quad_perm [1,0,3,2] (0xB1): i <- i ^ 1;  [2,3,0,1] (0x4E): i <- i ^ 2
//   DPP row_half_mirror (0x141): i <- 7 - i within 8; row_mirror (0x140):
//     i <- 15 - i within 16
//   ds_swizzle bitmask 0x401F: i <- i ^ 16 within 32
//   v_permlane32_swap(v, v): [0] = v[l mod 32], [1] = v[32 + l mod 32]
//   v_permlane16_swap(v, v): [0] = v[32 (l / 32) + l mod 16],
//     [1] = v[32 (l / 32) + 16 + l mod 16]
// The permlane swaps are VALU operations; ds_swizzle goes through the LDS
// unit and waits for it.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// value of lane l ^ 16
__device__ __forceinline__ uint32_t xor16_u32(uint32_t v) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (__lane_id() & 16) ? p[0] : p[1];
}
__device__ __forceinline__ double xor16_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = xor16_u32((uint32_t)b), hi = xor16_u32((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// value of lane l ^ 32, one 32-bit swap
__device__ __forceinline__ float xor32_f32(float v) {
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    const auto p = __builtin_amdgcn_permlane32_swap(b, b, false, false);
    return __builtin_bit_cast(float, __lane_id() < 32 ? p[1] : p[0]);
}
// value of lane l ^ 32
__device__ __forceinline__ double xor32_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)b, (uint32_t)b, false, false);
    const auto hi =
        __builtin_amdgcn_permlane32_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
    const bool low = __lane_id() < 32;
    const uint32_t l = low ? lo[1] : lo[0], h = low ? hi[1] : hi[0];
    return __builtin_bit_cast(double, ((uint64_t)h << 32) | l);
}

// lane l's value, broadcast (l wave-uniform)
__device__ __forceinline__ double readlane_lu(double v, int l) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float readlane_lu(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double xor32_lu(double v) { return xor32_f64(v); }
__device__ __forceinline__ float xor32_lu(float v) { return xor32_f32(v); }

// 1 / a for a pivot of the whole-wave elimination: the hardware reciprocal
// refined by Newton steps (two for double, one for float), within an ulp or
// two of the correctly rounded quotient, in 3-5 instructions instead of the
// IEEE division's ten.  The inverse is an approximation of the Jacobian's
// anyway (the Broyden iteration only needs a nearby matrix); a is never 0
// here (the pivot key rejected zeros).
__device__ __forceinline__ double pivot_recip(double a) {
    double r = __builtin_amdgcn_rcp(a);
    r = fma(r, fma(-a, r, 1.0), r);
    return fma(r, fma(-a, r, 1.0), r);
}
__device__ __forceinline__ float pivot_recip(float a) {
    const float r = __builtin_amdgcn_rcpf(a);
    return fmaf(r, fmaf(-a, r, 1.0f), r);
}

// ---------------------------------------------------------------------------
// group collectives (every lane of the group active)
// ---------------------------------------------------------------------------
template <int W>
struct Grp {
    int g;  // lane index in the group
    __device__ __forceinline__ double bcast(double v, int src) const {
        if constexpr (W == 1) return v;
        return __shfl(v, src, W);
    }
    // all-reduce sum; symmetric pairings, so every lane gets the same bits
    __device__ __forceinline__ double sum(double v) const {
        if constexpr (W == 32) {
            v += dpp_f64<0xB1>(v);
            v += dpp_f64<0x4E>(v);
            v += dpp_f64<0x141>(v);
            v += dpp_f64<0x140>(v);
            v += xor16_f64(v);
            return v;
        } else {
#pragma unroll
            for (int m = W / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, W);
            return v;
        }
    }
    // all-reduce maximum of unsigned keys (W == 32: one DPP max per stage)
    __device__ __forceinline__ uint32_t max_u32(uint32_t v) const {
        if constexpr (W == 32) {
            v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true));
            v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true));
            v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true));
            v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true));
            v = __builtin_elementwise_max(v, xor16_u32(v));
            return v;
        } else {
#pragma unroll
            for (int m = W / 2; m >= 1; m >>= 1)
                v = __builtin_elementwise_max(v, (uint32_t)__shfl_xor((int)v, m, W));
            return v;
        }
    }
    // all-reduce maximum (IEEE maxNum: a NaN operand loses)
    __device__ __forceinline__ double max(double v) const {
        if constexpr (W == 32) {
            v = fmax(v, dpp_f64<0xB1>(v));
            v = fmax(v, dpp_f64<0x4E>(v));
            v = fmax(v, dpp_f64<0x141>(v));
            v = fmax(v, dpp_f64<0x140>(v));
            v = fmax(v, xor16_f64(v));
            return v;
        } else {
#pragma unroll
            for (int m = W / 2; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, W));
            return v;
        }
    }
};

// maximum of unsigned keys over each 16-lane row (one DPP max per stage)
__device__ __forceinline__ uint32_t row_max_u32(uint32_t v) {
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true));
    return v;
}

// Pivot choice key: |a| with its 5 lowest mantissa bits replaced by 31 - g,
// so the maximum key names the row with the largest |a| and the lowest lane
// among equal ones (ties below 2^-47 relative resolve towards lower lanes).
// Rows not eligible (already pivots, NaN) get key 0.
__device__ __forceinline__ double pivot_key(double a, bool eligible, int g) {
    const double m = fabs(a);
    if (!eligible || !(m == m)) return 0.0;
    const uint64_t b = (__builtin_bit_cast(uint64_t, m) & ~(uint64_t)31) | (uint64_t)(31 - g);
    return __builtin_bit_cast(double, b);
}
__device__ __forceinline__ int pivot_lane(double key) {
    return 31 - (int)(__builtin_bit_cast(uint64_t, key) & 31);
}
__device__ __forceinline__ bool pivot_nonzero(double key) {
    return (__builtin_bit_cast(uint64_t, key) & ~(uint64_t)31) != 0;
}

// The same choice on 32-bit keys: the high word of |a| (sign, exponent and
// 20 mantissa bits) with its 5 lowest bits replaced by 31 - g, reduced by
// one DPP integer max per stage.  The largest |a| wins up to 2^-15
// relative (candidates that close are equally good pivots); ties go to the
// lowest lane.  |a| below 2^-1017 counts as zero (SingularMatrix), as do
// rows already used and NaN.
__device__ __forceinline__ uint32_t pivot_key32(double a, bool eligible, int g) {
    if (!eligible || !(a == a)) return 0u;
    const uint32_t hi = (uint32_t)(__builtin_bit_cast(uint64_t, a) >> 32) & 0x7fffffffu;
    return (hi & ~31u) | (uint32_t)(31 - g);
}
// The FP32 inverse's key straight from the float: |a|'s exponent and 23
// mantissa bits, the 5 lowest replaced by 31 - g (the largest |a| wins up
// to 2^-18 relative); |a| below 2^-144 counts as zero.
__device__ __forceinline__ uint32_t pivot_key32(float a, bool eligible, int g) {
    if (!eligible || !(a == a)) return 0u;
    const uint32_t m = __builtin_bit_cast(uint32_t, a) & 0x7fffffffu;
    return (m & ~31u) | (uint32_t)(31 - g);
}
__device__ __forceinline__ int pivot_lane32(uint32_t key) { return 31 - (int)(key & 31u); }
__device__ __forceinline__ bool pivot_nonzero32(uint32_t key) { return (key & ~31u) != 0u; }

// ---------------------------------------------------------------------------
// right-hand sides evaluated cooperatively: lane g returns f_g(t, y)
// ---------------------------------------------------------------------------
template <template <int> class R>
struct ScalarBdf {  // the 1-D reference-test right-hand sides
    static constexpr int D = 1, W = 1, P = R<1>::P;
    template <class G>
    __device__ static bool lane_eval(double t, double y, const double *p, const G &, double &f) {
        double yy[1] = {y}, ff[1];
        const bool ok = R<1>::eval(t, yy, p, ff);
        f = ff[0];
        return ok;
    }
};

// 1-D Brusselator reaction-diffusion, CELLS cells, periodic, dx = 1/CELLS;
// y = [u_0..u_{CELLS-1}, v_0..v_{CELLS-1}], params A, B, alpha.
template <int CELLS>
struct RhsBrusselatorG {
    static constexpr int D = 2 * CELLS, W = D, P = 3;
    template <class G>
    __device__ static bool lane_eval(double, double y, const double *p, const G &grp, double &f) {
        const int g = grp.g;
        const bool isu = g < CELLS;
        const int c = isu ? g : g - CELLS;
        const int base = isu ? 0 : CELLS;
        double left, right, other;
        if constexpr (CELLS == 16) {
            // a species' 16 cells fill one DPP row: periodic neighbours are
            // row rotations, the other species is the lane 16 apart
            left = dpp_f64<0x121>(y);
            right = dpp_f64<0x12F>(y);
            other = xor16_f64(y);
        } else {
            left = grp.bcast(y, base + (c + CELLS - 1) % CELLS);
            right = grp.bcast(y, base + (c + 1) % CELLS);
            other = grp.bcast(y, isu ? g + CELLS : g - CELLS);
        }
        const double u = isu ? y : other, v = isu ? other : y;
        const double A = p[0], B = p[1], alpha = p[2];
        const double dx = 1.0 / CELLS;
        const double cl = alpha / (dx * dx);
        const double uuv = u * u * v;
        const double lap = fma(-2.0, y, left) + right;
        // u: A + u^2 v - (B + 1) u + cl lap;  v: B u - u^2 v + cl lap
        const double react = isu ? fma(-(B + 1.0), u, A + uuv) : fma(B, u, -uuv);
        f = fma(cl, lap, react);
        return true;
    }
    // FLOPs of one full evaluation (all 2*CELLS components)
    static constexpr int FLOPS = CELLS * (6 + 4 + 4 + 5);

    // Column colouring for the finite-difference Jacobian (Curtis, Powell &
    // Reid 1974): row u_i depends on u_{i-1}, u_i, u_{i+1}, v_i and row v_i
    // on v_{i-1}, v_i, v_{i+1}, u_i.  Colour u_j with j mod 4 and v_j with
    // (j + 2) mod 4: the four members of every row's stencil then carry the
    // four colours, so one evaluation per colour (and sign) yields one
    // entry of every row (CELLS divisible by 4 keeps the periodic wrap
    // consistent).
    static constexpr int NC = 4;
    // the column of colour k in row g's stencil
    __device__ static int member(int g, int k) {
        const bool isu = g < CELLS;
        const int i = isu ? g : g - CELLS;
        const int d = (k - i) & 3;
        if (isu) {
            const int c[4] = {g, (i + 1) % CELLS, CELLS + i, (i + CELLS - 1) % CELLS};
            return c[d];
        }
        const int c[4] = {i, CELLS + (i + CELLS - 1) % CELLS, g, CELLS + (i + 1) % CELLS};
        return c[d];
    }
    // row g's stencil in increasing column order
    __device__ static void stencil(int g, int (&s)[4]) {
        const bool isu = g < CELLS;
        const int i = isu ? g : g - CELLS;
        const int base = isu ? 0 : CELLS;
        int a[4] = {base + (i + CELLS - 1) % CELLS, g, base + (i + 1) % CELLS,
                    isu ? CELLS + i : i};
        // sort four (a tiny network; compile-time unrolled)
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3 - x; ++y)
                if (a[y] > a[y + 1]) {
                    const int tmp = a[y];
                    a[y] = a[y + 1];
                    a[y + 1] = tmp;
                }
#pragma unroll
        for (int x = 0; x < 4; ++x) s[x] = a[x];
    }
    // f_g at the state the reference's column loop (bdf.rs:402-406) holds
    // while it evaluates column c: components j < c already restored
    // (`rest`, x + dt - 2 dt + dt), j > c original (`orig`), c itself
    // perturbed (`pert`: x + dt on half 0, x + dt - 2 dt on half 1).  Every
    // lane supplies its own three candidates; `own` receives the value of
    // the lane's own component in that state.  A c outside the stencil
    // gives the unperturbed state of that column.
    template <class G>
    __device__ static bool lane_eval_state(double t, double orig, double rest, double pert, int c,
                                           const double *p, const G &grp, double &f,
                                           double &own) {
        static_assert(CELLS == 16, "state evaluation uses the DPP row layout of 16 cells");
        const int g = grp.g;
        const bool isu = g < CELLS;
        const int i = isu ? g : g - CELLS;
        const int base = isu ? 0 : CELLS;
        const int jl = base + (i + CELLS - 1) % CELLS, jr = base + (i + 1) % CELLS;
        const int jo = isu ? g + CELLS : g - CELLS;
        auto pick = [&](int j, double o, double r, double q) {
            return j == c ? q : (j < c ? r : o);
        };
        own = pick(g, orig, rest, pert);
        const double left = pick(jl, dpp_f64<0x121>(orig), dpp_f64<0x121>(rest), dpp_f64<0x121>(pert));
        const double right =
            pick(jr, dpp_f64<0x12F>(orig), dpp_f64<0x12F>(rest), dpp_f64<0x12F>(pert));
        const double other = pick(jo, xor16_f64(orig), xor16_f64(rest), xor16_f64(pert));
        const double u = isu ? own : other, v = isu ? other : own;
        const double A = p[0], B = p[1], alpha = p[2];
        const double dx = 1.0 / CELLS;
        const double cl = alpha / (dx * dx);
        const double uuv = u * u * v;
        const double lap = fma(-2.0, own, left) + right;
        const double react = isu ? fma(-(B + 1.0), u, A + uuv) : fma(B, u, -uuv);
        f = fma(cl, lap, react);
        return true;
    }
};

// A small system given as a register right-hand side (bg_device.h: f from
// the whole state): every lane of the group gathers the D components,
// evaluates f and keeps its own component.  The group is the power of two
// W >= D; lanes g >= D are padding that hold 0 in every vector, 0 as f, and
// the identity row and column of the Jacobian, so they never affect a sum,
// a pivot or an update.
template <class R>
struct GatherBdf {
    static constexpr int D = R::D, P = R::P;
    static constexpr int W = D <= 1 ? 1 : D <= 2 ? 2 : D <= 4 ? 4 : D <= 8 ? 8 : D <= 16 ? 16 : 32;
    static_assert(D <= 16, "gathered systems take at most a 16-lane group");
    template <class G>
    __device__ static bool lane_eval(double t, double y, const double *p, const G &grp, double &f) {
        double yy[D], ff[D];
#pragma unroll
        for (int k = 0; k < D; ++k) yy[k] = grp.bcast(y, k);
        const bool ok = R::eval(t, yy, p, ff);
        f = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) f = grp.g == k ? ff[k] : f;
        return ok;
    }
};

// quirk register as compile-time data (bdf.rs:403-407, 568, 622)
template <bool AsBuilt>
struct BdfSem {
    static constexpr double fd_sign = AsBuilt ? 1.0 : -1.0;  // Q3
    static constexpr bool lower_own_hist = !AsBuilt;         // Q4
    static constexpr double theta = AsBuilt ? 0.0 : 1.0;     // Q5
    static constexpr bool restore_mul = !AsBuilt;            // Q6
};

// BDF coefficients (bdf.rs:646-672, 713-729)
template <int O>
struct BdfCoef;
template <>
struct BdfCoef<7> {
    __device__ static constexpr double hi(int k) {
        constexpr double h = 147.0;
        constexpr double c[7] = {60.0 / h, -360.0 / h, 450.0 / h, -400.0 / h,
                                 225.0 / h, -72.0 / h, 10.0 / h};
        return c[k];
    }
    __device__ static constexpr double lo(int k) {
        constexpr double l = 137.0;
        constexpr double c[7] = {60.0 / l, -300.0 / l, 300.0 / l, -200.0 / l,
                                 75.0 / l, -12.0 / l, 0.0};
        return c[k];
    }
};
template <>
struct BdfCoef<3> {
    __device__ static constexpr double hi(int k) {
        constexpr double c[3] = {2.0 / 3.0, -4.0 / 3.0, 1.0 / 3.0};
        return c[k];
    }
    __device__ static constexpr double lo(int k) {
        constexpr double c[3] = {1.0, -1.0, 0.0};
        return c[k];
    }
};

enum : int { BDF_YIELD = -1, BDF_REDO = -2, BDF_DONE = -3 };

// ---------------------------------------------------------------------------
// one group's BDF solver state (per lane: its component / its row)
// ---------------------------------------------------------------------------
// SPLIT = 1: a group of W = D lanes per instance; lane g owns component g and
// row g of the Jacobian inverse.  SPLIT = 2: the whole wave per instance
// (2 W = 64 lanes): lane (h, g) = h W + g owns component g (both halves hold
// the same component values, computed redundantly) and the half row
// Jinv[g][h W/2 .. h W/2 + W/2 - 1].  The split halves the registers for the
// inverse, runs the FD Jacobian's +dt and -dt evaluations on the two halves
// at once, and leaves no second instance in the wave to diverge from.
template <class RHS, int O, class Sem, class LuT, int SPLIT = 1>
struct BdfLane {
    // D components in a group of W lanes (W = D except for padded gathered
    // systems); SPLIT = 2 only for D = W = 32
    static constexpr int D = RHS::D, W = RHS::W, WJ = W / SPLIT;
    using C = BdfCoef<O>;
    Grp<W> grp;
    int h = 0;    // half of the wave (SPLIT = 2)
    double *lds;  // this group's W x (W+1) transpose buffer, then 2 scalars
    float *prow_f = nullptr;  // SPLIT = 2, FP32 inverse: the pivot row

    // The elimination's pivot row (W entries) in the inverse's arithmetic
    // type: FP64 uses the group's LDS buffer (its row 0, free during the
    // elimination), FP32 a float buffer of its own.  1 / pivot goes from the
    // pivot lane by readlane.
    __device__ __forceinline__ LuT *pivot_row() const {
        if constexpr (sizeof(LuT) == sizeof(double)) return lds;
        else return prow_f;
    }
    // instance scalars (replicated across the group)
    double t, dt, end, tol, dt_min, dt_max;
    int yield_memory;
    bool hist_full;
    double hist_t[O];
    // per-lane component data
    double y, save;
    double hist[O];  // prev_values front..back (own component)
    double jinv[WJ];  // row g of the Jacobian inverse (SPLIT = 2: half row h)
    double p[RHS::P > 0 ? RHS::P : 1];
    // counters
    uint32_t n_newton, n_lu;

    __device__ __forceinline__ bool f(double tt, double yy, double &out) const {
        return RHS::lane_eval(tt, yy, p, grp, out);
    }

    // hist push_back + pop_front on a full deque
    __device__ __forceinline__ void hist_shift(double tt, double yy) {
#pragma unroll
        for (int k = 0; k < O - 1; ++k) {
            hist[k] = hist[k + 1];
            hist_t[k] = hist_t[k + 1];
        }
        hist[O - 1] = yy;
        hist_t[O - 1] = tt;
    }

    // bdf.rs:346-387: classical RK4, `iterations` steps; with fill = true the
    // deque is rebuilt from the states after steps 1..O (O == iterations).
    template <int IT, bool FILL>
    __device__ __forceinline__ int runge_kutta() {
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            double k1, k2, k3, k4;
            if (!f(t, y, k1)) return BG_USER_ERROR;
            k1 *= dt;
            if (!f(t + 0.5 * dt, y + k1 * 0.5, k2)) return BG_USER_ERROR;
            k2 *= dt;
            if (!f(t + 0.5 * dt, y + k2 * 0.5, k3)) return BG_USER_ERROR;
            k3 *= dt;
            if (!f(t + dt, y + k3, k4)) return BG_USER_ERROR;
            k4 *= dt;
            if (FILL && i != 0) {
                hist[i - 1] = y;
                hist_t[i - 1] = t;
            }
            y += (k1 + k2 * 2.0 + k3 * 2.0 + k4) * (1.0 / 6.0);
            t += dt;
        }
        if (FILL) {
            hist[IT - 1] = y;
            hist_t[IT - 1] = t;
            hist_full = true;
        }
        return BG_OK;
    }

    // residual closures bdf.rs:548-575 (lane g's component), given
    // fx = f(t_n, x)_g
    template <bool LOWER>
    __device__ __forceinline__ double combine(double fx, double x) const {
        constexpr bool own = LOWER && Sem::lower_own_hist;
        double g = (-fx) * dt * (LOWER ? C::lo(0) : C::hi(0));
#pragma unroll
        for (int ind = 1; ind < O; ++ind) {
            const double c = own ? C::lo(ind) : C::hi(ind);
            if (c != 0.0) g = fma(hist[O - ind], c, g);
        }
        return g + x;
    }

    template <bool LOWER>
    __device__ __forceinline__ bool residual(double tt, double x, double &g) const {
        double fx;
        if (!f(tt, x, fx)) return false;
        g = combine<LOWER>(fx, x);
        return true;
    }

    // y_g = -(Jinv x)_g with x broadcast through LDS
    __device__ __forceinline__ double neg_jinv_times(double x) {
        lds[grp.g] = x;
        __builtin_amdgcn_wave_barrier();
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < WJ; ++j) acc = fma(jinv[j], lds[h * WJ + j], acc);
        if constexpr (SPLIT == 2) acc += xor32_f64(acc);
        __builtin_amdgcn_wave_barrier();
        return -acc;
    }

    // All-reduce sum of the component values.  SPLIT = 2: both halves hold
    // the same values, so half 0's two 16-lane row sums, added once,
    // replace the cross-row swap (the same bits: r0 + r1 = r1 + r0).
    __device__ __forceinline__ double comp_sum(double v) const {
        if constexpr (SPLIT == 2) {
            v += dpp_f64<0xB1>(v);
            v += dpp_f64<0x4E>(v);
            v += dpp_f64<0x141>(v);
            v += dpp_f64<0x140>(v);
            return readlane_lu(v, 0) + readlane_lu(v, 16);
        } else {
            return grp.sum(v);
        }
    }

    // In-place Gauss-Jordan inversion (exchange steps) of the group matrix
    // whose row g is staged in row g of the LDS buffer, with partial pivoting
    // on rows that are not swapped: column k pivots on the unused row p with
    // the largest |a(p,k)| (pivot_key: lowest lane on ties).  Afterwards lane
    // p holds row col(p) of the inverse with its columns permuted by col(),
    // which one LDS scatter/gather undoes.  Arithmetic in LuT; returns false
    // on a zero (or NaN) pivot.
    __device__ __forceinline__ bool invert() {
        if constexpr (SPLIT == 2) return invert_split();
        else return invert_group();
    }

    // SPLIT = 2: the same exchange steps on half rows.  Column k lives in
    // half k / WJ; the other half fetches a(g, k) across the halves.
    // Normalisation is deferred: the pivot row is published unscaled with
    // 1 / pivot beside it (LDS slot W, the padding column of row 0, unused
    // until the final scatter), the other rows eliminate with
    // f = a(g, k) / pivot, and each row is scaled by its own 1 / pivot once
    // at the end.  A row that has been a pivot row holds pivot times its
    // normalised values, so later updates stay consistent; this saves the
    // per-step scaling of the pivot row.
    __device__ __forceinline__ bool invert_split() {
        LuT b[WJ];
#pragma unroll
        for (int j = 0; j < WJ; ++j) b[j] = (LuT)lds[grp.g * (W + 1) + h * WJ + j];
        __builtin_amdgcn_wave_barrier();
        bool used = false;
        // the smallest key so far: a running minimum, so no column's key
        // stays live until the end of the elimination
        uint32_t key_min = ~0u;
        int my_col = 0;
        LuT my_inv = (LuT)1;
        LuT *const prow = pivot_row();
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int hk = k / WJ, kk = k % WJ;
            const LuT mk = b[kk];
            // half hk holds column k in its own lanes: its reduction is the
            // key (the other half reduces another column, unused), and the
            // lane swap for the other half's copy stays off the search
            // (the maxima of its two 16-lane rows, combined in scalars)
            const uint32_t rm = row_max_u32(pivot_key32(mk, !used, grp.g));
            const uint32_t key = __builtin_elementwise_max(
                (uint32_t)__builtin_amdgcn_readlane((int)rm, hk * W),
                (uint32_t)__builtin_amdgcn_readlane((int)rm, hk * W + 16));
            const LuT other = xor32_lu(mk);
            const LuT agk = (h == hk) ? mk : other;
            key_min = __builtin_elementwise_min(key_min, key);
            const int pl = pivot_lane32(key);
            const bool piv = (pl == grp.g);
            if (piv) {
                used = true;
                my_col = k;
#pragma unroll
                for (int j = 0; j < WJ; ++j) prow[h * WJ + j] = b[j];
                my_inv = pivot_recip(agk);
            }
            __builtin_amdgcn_wave_barrier();
            // 1 / pivot straight from the pivot lane (half 0), not via LDS
            const LuT inv = readlane_lu(my_inv, pl);
            // Every lane updates, so the next column's pivot search can
            // overlap this step's FMAs.  The pivot row's lanes read back
            // their own row and take factor +0: fma(+0, b, b) = b for every
            // finite b, zeros' signs included (an infinite entry turns into
            // NaN instead of staying infinite: the inverse is not finite
            // either way).
            const LuT f = agk * inv;
            const LuT nf = piv ? (LuT)0 : -f;
#pragma unroll
            for (int j = 0; j < WJ; ++j) b[j] = fma(nf, prow[h * WJ + j], b[j]);
            // column k of the inverse: pivot * (1 / pivot) on the pivot row,
            // -a(g, k) / pivot elsewhere
            if (h == hk) b[kk] = piv ? (LuT)1 : -f;
            __builtin_amdgcn_wave_barrier();
        }
        const bool ok = pivot_nonzero32(key_min);
#pragma unroll
        for (int j = 0; j < WJ; ++j) b[j] *= my_inv;
#pragma unroll
        for (int c = 0; c < WJ; ++c) lds[my_col * (W + 1) + h * WJ + c] = (double)b[c];
        lds[grp.g * (W + 1) + W] = (double)my_col;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < WJ; ++q) {
            int cq = (int)lds[(h * WJ + q) * (W + 1) + W];
            cq = cq < 0 ? 0 : (cq >= W ? W - 1 : cq);
            jinv[q] = lds[grp.g * (W + 1) + cq];
        }
        __builtin_amdgcn_wave_barrier();
        return ok;
    }

    __device__ __forceinline__ bool invert_group() {
        LuT b[W];
#pragma unroll
        for (int j = 0; j < W; ++j) b[j] = (LuT)lds[grp.g * (W + 1) + j];
        __builtin_amdgcn_wave_barrier();
        bool used = false, ok = true;
        int my_col = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const double key = grp.max(pivot_key((double)b[k], !used, grp.g));
            if (!pivot_nonzero(key)) ok = false;
            const bool piv = (pivot_lane(key) == grp.g);
            if (piv) {
                used = true;
                my_col = k;
                const LuT ip = (LuT)1 / b[k];
#pragma unroll
                for (int j = 0; j < W; ++j)
                    if (j != k) b[j] *= ip;
                b[k] = ip;
#pragma unroll
                for (int j = 0; j < W; ++j) lds[j] = (double)b[j];
            }
            __builtin_amdgcn_wave_barrier();
            if (!piv) {
                const LuT f = b[k];
#pragma unroll
                for (int j = 0; j < W; ++j)
                    if (j != k) b[j] = fma(-f, (LuT)lds[j], b[j]);
                b[k] = -f * (LuT)lds[k];
            }
            __builtin_amdgcn_wave_barrier();
        }
        // lane p: row col(p) of the inverse, column c belongs to inverse
        // column q with col(q) = c.  Scatter rows by col(), publish col() in
        // the padding column, gather with the column permutation.
#pragma unroll
        for (int c = 0; c < W; ++c) lds[my_col * (W + 1) + c] = (double)b[c];
        lds[grp.g * (W + 1) + W] = (double)my_col;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int cq = (int)lds[q * (W + 1) + W];
            cq = cq < 0 ? 0 : (cq >= W ? W - 1 : cq);
            jinv[q] = lds[grp.g * (W + 1) + cq];
        }
        __builtin_amdgcn_wave_barrier();
        return ok;
    }

    // bdf.rs:390-411 for the whole-wave layout (SPLIT = 2): the
    // finite-difference Jacobian by column colouring.  The reference's loop
    // perturbs one column at a time (x_c + dt, then x_c + dt - 2 dt, then
    // restores x_c as x_c + dt - 2 dt + dt, which may differ from x_c by a
    // rounding), so column c is evaluated at a state whose components
    // before c are restored and after c original.  Entry (g, c) depends only
    // on row g's stencil, so one evaluation per colour (half 0 at + dt,
    // half 1 at - dt) gives every row its entry of that colour, each lane at
    // the state its own column c would see: the same inputs, hence the same
    // bits, as the column loop.  Entries outside the stencil see identical
    // inputs above and below: the central difference (intended) makes them
    // exactly 0; the sum (as built, Q3) makes them (G + G) / (2 dt) at the
    // unperturbed state of their column, which depends only on how many
    // stencil members precede the column (at most five states).  Lane
    // (h, g) writes the entries of its half row h into row g of the LDS
    // buffer.  Returns the guess after the loop (every component restored).
    template <bool LOWER>
    __device__ __forceinline__ bool fd_coloured(double tt, double denom, double &guess) {
        const double orig = guess;
        const double xp = orig + dt, xm = xp - 2.0 * dt;
        const double rest = xm + dt;  // the reference's restore (bdf.rs:405)
        const double pert = h == 0 ? xp : xm;
        const int g = grp.g;
        double *row = lds + g * (W + 1);
        if constexpr (Sem::fd_sign > 0.0) {
            // as built: the dense part, from the unperturbed states
            int s[4];
            RHS::stencil(g, s);
            double Gm[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                // a column with exactly m stencil members before it, if the
                // row has one (otherwise the value is not used)
                const int rep = m == 0 ? 0 : s[m - 1] + 1;
                double fx, own;
                if (!RHS::lane_eval_state(tt, orig, rest, pert, rep, p, grp, fx, own)) return false;
                Gm[m] = combine<LOWER>(fx, own);
            }
#pragma unroll
            for (int j = 0; j < WJ; ++j) {
                const int c = h * WJ + j;
                const int m = (s[0] < c) + (s[1] < c) + (s[2] < c) + (s[3] < c);
                double v = Gm[0];
#pragma unroll
                for (int q = 1; q < 5; ++q) v = m == q ? Gm[q] : v;
                row[c] = (v + Sem::fd_sign * v) * denom;
            }
        } else {
#pragma unroll
            for (int j = 0; j < WJ; ++j) row[h * WJ + j] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < RHS::NC; ++k) {
            const int c = RHS::member(g, k);
            double fx, own;
            if (!RHS::lane_eval_state(tt, orig, rest, pert, c, p, grp, fx, own)) return false;
            const double r = combine<LOWER>(fx, own);
            const double other = xor32_f64(r);
            const double above = h == 0 ? r : other, below = h == 0 ? other : r;
            if (c / WJ == h) row[c] = (above + Sem::fd_sign * below) * denom;
        }
        guess = rest;
        return true;
    }

    // bdf.rs:390-411: the finite-difference Jacobian into the LDS rows,
    // starting from guess = y; leaves the guess as the reference's column
    // loop does.  False on a UserError.
    template <bool LOWER>
    __device__ __forceinline__ bool fd_jacobian(double tt, double denom, double &guess) {
        guess = y;
        if constexpr (SPLIT == 2) {
            if (!fd_coloured<LOWER>(tt, denom, guess)) return false;
        } else {
        // columns (G(x + dt e_c) +/- G(x - dt e_c)) / (2 dt); lane g stores
        // J(g, c) into row g of the group's LDS buffer
#pragma unroll 1
        for (int c = 0; c < W; ++c) {
            if (c >= D) {  // padding column: e_c
                lds[grp.g * (W + 1) + c] = grp.g == c ? 1.0 : 0.0;
                continue;
            }
            const bool mine = (grp.g == c);
            if (mine) guess += dt;
            double above, below;
            if (!residual<LOWER>(tt, guess, above)) return false;
            if (mine) guess -= 2.0 * dt;
            if (!residual<LOWER>(tt, guess, below)) return false;
            if (mine) guess += dt;
            lds[grp.g * (W + 1) + c] = (above + Sem::fd_sign * below) * denom;
        }
        }
        __builtin_amdgcn_wave_barrier();
        return true;
    }

    // ---- the rest of the inverse cascade (bdf.rs:429-444) -----------------
    // nalgebra tries partial-pivot LU, then full-pivot LU, then QR, and
    // reports SingularMatrix only when all three meet an exact zero pivot.
    // The fast path is the partial-pivot elimination above; when it finds
    // no pivot, these two run on the same matrix (rebuilt in the LDS rows by
    // fd_jacobian).  They are rare, so they work in LDS with rolled loops
    // (few registers) and keep their per-lane temporaries in private memory.
    // Both halves of a whole-wave group compute the same values; every lane
    // ends holding its part of row g of the inverse in jinv.
    __device__ __forceinline__ double &mat(int r, int c) const { return lds[r * (W + 1) + c]; }

    // row g of the inverse, given as the matrix M whose row qinv(g) holds
    // row g with column j at position col(j); col(r) is in mat(r, W).
    __device__ __forceinline__ void gather_permuted_row() {
        const int g = grp.g;
        int src = 0;
#pragma unroll 1
        for (int r = 0; r < W; ++r) src = ((int)mat(r, W) == g) ? r : src;
        double row[WJ];
#pragma unroll
        for (int q = 0; q < WJ; ++q) {
            int cq = (int)mat(h * WJ + q, W);
            cq = cq < 0 ? 0 : (cq >= W ? W - 1 : cq);
            row[q] = mat(src, cq);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < WJ; ++q) jinv[q] = row[q];
    }

    // Full-pivot Gauss-Jordan (exchange steps): step k pivots on the largest
    // |a| over the unused rows and columns (first column, then lowest row,
    // on ties) and fails on an exact zero, like FullPivLU's try_inverse.
    __device__ __forceinline__ bool invert_full_pivot() {
        const int g = grp.g;
        bool used = false;
        uint32_t ucols = 0;  // columns already pivoted (the same in every lane)
        int my_col = -1;
#pragma unroll 1
        for (int k = 0; k < W; ++k) {
            double best = -1.0;
            int bc = 0;
            if (!used) {
#pragma unroll 1
                for (int c = 0; c < W; ++c) {
                    const double v = fabs(mat(g, c));
                    if (!((ucols >> c) & 1u) && v > best) {
                        best = v;
                        bc = c;
                    }
                }
            }
            const double gmax = grp.max(best);
            if (!(gmax > 0.0)) return false;
            const uint32_t key = grp.max_u32((!used && best == gmax) ? (uint32_t)(63 - g) : 0u);
            const int p = 63 - (int)key;
            const int q = (int)grp.bcast((double)bc, p);
            const double ip = 1.0 / mat(p, q);
            __builtin_amdgcn_wave_barrier();
            if (g == p) {
#pragma unroll 1
                for (int j = 0; j < W; ++j) mat(p, j) = j == q ? ip : mat(p, j) * ip;
                used = true;
                my_col = q;
            }
            __builtin_amdgcn_wave_barrier();
            if (g != p) {
                const double fq = mat(g, q);
#pragma unroll 1
                for (int j = 0; j < W; ++j)
                    if (j != q) mat(g, j) = fma(-fq, mat(p, j), mat(g, j));
                mat(g, q) = -fq * ip;
            }
            ucols |= 1u << q;
            __builtin_amdgcn_wave_barrier();
        }
        mat(g, W) = (double)my_col;
        __builtin_amdgcn_wave_barrier();
        gather_permuted_row();
        return true;
    }

    // Householder QR (reflections chosen away from cancellation), failing on
    // an exact zero on R's diagonal like QR's try_inverse.  Reflection k is
    // kept below the diagonal of column k, scaled to a unit leading entry,
    // with its factor in mat(k, W).  Row g of the inverse R^-1 Q^T is
    // (Q w)^T with R^T w = e_g, solved by each lane on its own.
    __device__ __forceinline__ bool invert_qr() {
        const int g = grp.g;
#pragma unroll 1
        for (int k = 0; k < W; ++k) {
            const double x = g >= k ? mat(g, k) : 0.0;
            const double nrm = sqrt(grp.sum(x * x));
            const double akk = mat(k, k);
            const double alpha = akk >= 0.0 ? -nrm : nrm;
            const double vk = akk - alpha;
            double v = g > k ? x : (g == k ? vk : 0.0);
            const double vv = grp.sum(v * v);
            const bool skip = nrm == 0.0 || vv == 0.0;
            const double beta = skip ? 0.0 : 2.0 / vv;
#pragma unroll 1
            for (int c = k; c < W; ++c) {
                const double dot = grp.sum(v * mat(g, c));
                __builtin_amdgcn_wave_barrier();
                if (g >= k) mat(g, c) = fma(-(beta * dot), v, mat(g, c));
                __builtin_amdgcn_wave_barrier();
            }
            // unit leading entry: v / vk, factor beta vk^2
            const double sv = skip ? 0.0 : v / vk;
            if (g > k) mat(g, k) = sv;
            if (g == k) mat(k, W) = skip ? 0.0 : beta * vk * vk;
            __builtin_amdgcn_wave_barrier();
        }
        bool ok = true;
#pragma unroll 1
        for (int k = 0; k < W; ++k) ok = ok && mat(k, k) != 0.0;
        if (!ok) return false;
        double w[W];
#pragma unroll 1
        for (int i = 0; i < W; ++i) {  // R^T w = e_g, forward
            double s = i == g ? 1.0 : 0.0;
#pragma unroll 1
            for (int l = 0; l < i; ++l) s = fma(-mat(l, i), w[l], s);
            w[i] = s / mat(i, i);
        }
#pragma unroll 1
        for (int k = W - 1; k >= 0; --k) {  // w <- H_k w, last reflection first
            const double bk = mat(k, W);
            double dot = w[k];
#pragma unroll 1
            for (int i = k + 1; i < W; ++i) dot = fma(mat(i, k), w[i], dot);
            const double sc = bk * dot;
            w[k] -= sc;
#pragma unroll 1
            for (int i = k + 1; i < W; ++i) w[i] = fma(-sc, mat(i, k), w[i]);
        }
#pragma unroll
        for (int q = 0; q < WJ; ++q) jinv[q] = w[h * WJ + q];
        __builtin_amdgcn_wave_barrier();
        return true;
    }

    // bdf.rs:414-475; on success `guess` holds the corrected value.
    template <bool LOWER>
    __device__ __forceinline__ int secant(double &guess) {
        const double tt = t + Sem::theta * dt;
        const double denom = 1.0 / (2.0 * dt);
        double F;
        guess = y;
        double fx0;
        if (!f(tt, guess, fx0)) return BG_USER_ERROR;
        F = combine<LOWER>(fx0, guess);
        if (!fd_jacobian<LOWER>(tt, denom, guess)) return BG_USER_ERROR;
        ++n_lu;
        if (!invert()) {
            // bdf.rs:432-441: full-pivot LU, then QR, each on the Jacobian
            // rebuilt from the same evaluations
            if (!fd_jacobian<LOWER>(tt, denom, guess)) return BG_USER_ERROR;
            if (!invert_full_pivot()) {
                if (!fd_jacobian<LOWER>(tt, denom, guess)) return BG_USER_ERROR;
                if (!invert_qr()) return BG_SINGULAR_MATRIX;
            }
        }
        double s = neg_jinv_times(F);
        guess += s;
#pragma unroll 1
        for (int n = 2; n < 1000; ++n) {
            const double Fl = F;
            if (!residual<LOWER>(tt, guess, F)) return BG_USER_ERROR;
            ++n_newton;
            const double adj = neg_jinv_times(F - Fl);
            // p = (-s)^T adj ; u = s^T Jinv (column sums through LDS)
            const double pp = comp_sum(adj * (-s));
#pragma unroll
            for (int j = 0; j < WJ; ++j) lds[grp.g * (W + 1) + h * WJ + j] = jinv[j] * s;
            __builtin_amdgcn_wave_barrier();
            double u = 0.0;
            if constexpr (SPLIT == 2) {
                // column g: rows of half h, then the other half's partial sum
#pragma unroll
                for (int k = 0; k < WJ; ++k) u += lds[(h * WJ + k) * (W + 1) + grp.g];
                u += xor32_f64(u);
            } else {
#pragma unroll
                for (int k = 0; k < W; ++k) u += lds[k * (W + 1) + grp.g];
            }
            __builtin_amdgcn_wave_barrier();
            lds[grp.g] = u;
            __builtin_amdgcn_wave_barrier();
            const double w = (s + adj) / pp;
#pragma unroll
            for (int j = 0; j < WJ; ++j) jinv[j] = fma(lds[h * WJ + j], w, jinv[j]);
            __builtin_amdgcn_wave_barrier();
            // s = -Jinv_new F (bdf.rs:464-465) without a second product
            // through LDS: with Jinv_new = Jinv + w u^T and Jinv F = -adj - s
            // (the previous s was -Jinv F_l), -Jinv_new F = adj + s - w (u . F);
            // equal in exact arithmetic, other roundings
            const double uf = comp_sum(u * F);
            s = fma(-w, uf, adj + s);
            guess += s;
            const double nrm = sqrt(comp_sum(s * s));
            if (nrm <= tol) return BG_OK;
        }
        return BG_MAXIMUM_ITERATIONS_EXCEEDED;
    }

    // One call of BDFSolver::step (bdf.rs:495-634).  Returns BDF_YIELD,
    // BDF_REDO, BDF_DONE or an IVPError code.  On a yield, (*yt, *yv) is the
    // yielded value (own component).
    __device__ __forceinline__ int step(double *yt, double *yv, bool *rejected) {
        *rejected = false;
        if (yield_memory > 0 && yield_memory <= O) {  // bdf.rs:500-512
            const int item = O - yield_memory;
            yield_memory -= 1;
            if (yield_memory == 0) yield_memory = O + 2;
            double vt = hist_t[0], vy = hist[0];
#pragma unroll
            for (int k = 1; k < O; ++k)
                if (item == k) {
                    vt = hist_t[k];
                    vy = hist[k];
                }
            *yt = vt;
            *yv = vy;
            return BDF_YIELD;
        }
        if (yield_memory == O + 2) {  // bdf.rs:519-525
            yield_memory = 0;
            hist_shift(t, y);
            *yt = t;
            *yv = y;
            return BDF_YIELD;
        }
        if (t >= end) return BDF_DONE;
        if (t + dt >= end) {  // bdf.rs:531-535
            dt = end - t;
            const int rc = runge_kutta<1, false>();
            if (rc) return rc;
            *yt = t;
            *yv = y;
            return BDF_YIELD;
        }
        if (!hist_full) {  // bdf.rs:537-546
            save = y;
            if (t + dt * (double)O >= end) dt = (end - t) / (double)O;
            const int rc = runge_kutta<O, true>();
            if (rc) return rc;
            yield_memory = O + 1;
            return BDF_REDO;
        }
        double hi, lo;
        int rc = secant<false>(hi);
        if (rc) return rc;
        rc = secant<true>(lo);
        if (rc) return rc;
        const double dlt = hi - lo;
        const double error = sqrt(grp.sum(dlt * dlt));
        if (error <= tol) {  // bdf.rs:583-614
            y = hi;
            t += dt;
            if (yield_memory == O + 1) {
                yield_memory -= 1;
                return BDF_REDO;
            }
            hist_shift(t, y);
            if (error < 0.1 * tol) {
                dt *= 2.0;
                if (dt > dt_max) dt = dt_max;
                hist_full = false;
            }
            *yt = t;
            *yv = y;
            return BDF_YIELD;
        }
        *rejected = true;
        if (yield_memory == O + 1) {  // bdf.rs:620-624
            if (Sem::restore_mul)
                t -= dt * (double)O;
            else
                t -= dt - (double)O;  // Q6
            y = save;
        }
        dt *= 0.5;
        if (dt < dt_min) return BG_MINIMUM_TIME_DELTA_EXCEEDED;
        hist_full = false;
        return BDF_REDO;
    }

    __device__ __forceinline__ void load(const bg_ensemble_desc &d, uint64_t i) {
        const uint64_t n = d.n;
        y = grp.g < D ? d.y0[(uint64_t)grp.g * n + i] : 0.0;
#pragma unroll
        for (int k = 0; k < RHS::P; ++k) p[k] = d.params[(uint64_t)k * n + i];
        t = d.t0 ? d.t0[i] : d.t0_s;
        end = d.t_end ? d.t_end[i] : d.t_end_s;
        tol = d.tol ? d.tol[i] : d.tol_s;
        dt_min = d.dt_min ? d.dt_min[i] : d.dt_min_s;
        dt_max = d.dt_max ? d.dt_max[i] : d.dt_max_s;
        dt = (dt_max + dt_min) * 0.5;  // bdf.rs:302
        yield_memory = 0;
        hist_full = false;
        save = y;
        n_newton = n_lu = 0;
#pragma unroll
        for (int k = 0; k < O; ++k) {
            hist[k] = 0.0;
            hist_t[k] = 0.0;
        }
    }
};

// ---------------------------------------------------------------------------
// persistent ensemble kernel: one instance per group, group refill
// ---------------------------------------------------------------------------
// Four waves per SIMD (the LDS limit): the solver is latency-bound (an LDS
// round trip on every elimination step), and waves hide each other's
// waits.  On the 32-component Brusselator, 4 waves with 101 VGPRs spilled
// measured 392 ms against 395 ms for 3 waves with 43 spilled (FP32 inverse:
// 240 against 248 ms); 3 had beaten 2 by 20 % before.
template <class RHS, int O, class Sem, class LuT, int SPLIT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
    bdf_ensemble_kernel(const bg_ensemble_desc *dv, uint32_t *__restrict__ counter) {
    constexpr int W = RHS::W;
    constexpr int GW = W * SPLIT;  // lanes per instance
    constexpr int GPW = 64 / GW;   // groups per wave
    __shared__ double lds_all[GPW][W * (W + 1) + 2];
    const int lane = __lane_id();
    const int grp_id = lane / GW;
    BdfLane<RHS, O, Sem, LuT, SPLIT> L;
    L.grp.g = lane % W;
    L.h = (lane % GW) / W;
    L.lds = lds_all[grp_id];
    // one wave per block; SPLIT = 2 holds one instance per wave
    if constexpr (SPLIT == 2 && sizeof(LuT) == sizeof(float)) {
        __shared__ float prow_all[W];
        L.prow_f = prow_all;
    }
    const uint64_t max_att = dv->max_attempts;
    const uint32_t cap =
        (uint32_t)((max_att && max_att < kBdfDefaultMaxSteps) ? max_att : kBdfDefaultMaxSteps);
    int64_t inst = -1;  // -1 need work, -2 exhausted
    for (;;) {
        // group leaders that need work
        const bool leader = (lane % GW == 0);
        const uint64_t need = __ballot(leader && inst == -1);
        if (need != 0) {
            const int first = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == first) base = atomicAdd(counter, (uint32_t)__popcll(need));
            base = __shfl(base, first);
            // rank of my group's leader among needing leaders
            const uint64_t below = need & ((1ull << (grp_id * GW)) - 1ull);
            const bool my_need = (need >> (grp_id * GW)) & 1ull;
            if (my_need) {
                const uint64_t k = (uint64_t)base + (uint64_t)__popcll(below);
                const bg_ensemble_desc di = desc_inputs(dv);
                if (k < di.n) {
                    inst = di.order ? (int64_t)di.order[k] : (int64_t)k;
                    L.load(di, (uint64_t)inst);
                } else {
                    inst = -2;
                }
            }
        }
        if (__ballot(inst >= 0) == 0) break;
        if (inst >= 0) {
            uint32_t steps = 0, nacc = 0, nrej = 0;
            int st;
            for (;;) {
                if (steps >= cap) {
                    st = BG_MAXIMUM_ITERATIONS_EXCEEDED;
                    break;
                }
                ++steps;
                double yt, yv;
                bool rej;
                const int r = L.step(&yt, &yv, &rej);
                if (rej) ++nrej;
                if (r == BDF_YIELD) {
                    ++nacc;
                } else if (r == BDF_DONE) {
                    st = BG_OK;
                    break;
                } else if (r != BDF_REDO) {
                    st = r;
                    break;
                }
            }
            const bg_ensemble_desc d = desc_outputs(dv);
            const uint64_t n = d.n, i = (uint64_t)inst;
            if (d.y_out && L.h == 0 && L.grp.g < RHS::D) d.y_out[(uint64_t)L.grp.g * n + i] = L.y;
            if (leader) {
                if (d.t_out) d.t_out[i] = L.t;
                if (d.status) d.status[i] = st;
                if (d.counts) {
                    bg_counts c;
                    c.n_accepted = nacc;
                    c.n_rejected = nrej;
                    d.counts[i] = c;
                }
                if (d.work) {
                    bg_counts c;
                    c.n_accepted = L.n_newton;
                    c.n_rejected = L.n_lu;
                    d.work[i] = c;
                }
            }
            inst = -1;
        }
    }
}

// Single-instance path production (one group): every yield is written into
// the path buffer; resumable when the buffer is full.
template <class RHS, int O, class Sem, class LuT>
__global__ void __launch_bounds__(64)
    bdf_path_kernel(BdfPathState *__restrict__ st, double *__restrict__ path_t,
                    double *__restrict__ path_y, uint32_t cap) {
    constexpr int D = RHS::D, W = RHS::W;
    __shared__ double lds[W * (W + 1) + 2];
    const int lane = __lane_id();
    if (lane >= W) return;
    BdfLane<RHS, O, Sem, LuT> L;
    L.grp.g = lane;
    L.lds = lds;
    L.t = st->t;
    L.dt = st->dt;
    L.end = st->end;
    L.tol = st->tol;
    L.dt_min = st->dt_min;
    L.dt_max = st->dt_max;
    L.yield_memory = st->yield_memory;
    L.hist_full = st->hist_full != 0;
    L.y = st->y[lane];
    L.save = st->save[lane];
#pragma unroll
    for (int k = 0; k < O; ++k) {
        L.hist[k] = st->hist[k][lane];
        L.hist_t[k] = st->hist_t[k];
    }
#pragma unroll
    for (int k = 0; k < RHS::P; ++k) L.p[k] = st->p[k];
    L.n_newton = st->n_newton;
    L.n_lu = st->n_lu;
    uint32_t count = 0;
    int status = BG_OK, finished = st->finished;
    uint64_t steps = st->steps;
    uint32_t nacc = st->n_accepted, nrej = st->n_rejected;
    while (!finished && count < cap) {
        if (steps >= st->max_steps) {
            status = BG_MAXIMUM_ITERATIONS_EXCEEDED;
            finished = 1;
            break;
        }
        ++steps;
        double yt, yv;
        bool rej;
        const int r = L.step(&yt, &yv, &rej);
        if (rej) ++nrej;
        if (r == BDF_YIELD) {
            ++nacc;
            if (lane == 0) path_t[count] = yt;
            if (lane < D) path_y[(uint64_t)count * D + lane] = yv;
            ++count;
        } else if (r == BDF_DONE) {
            finished = 1;
        } else if (r != BDF_REDO) {
            status = r;
            finished = 1;
        }
    }
    st->y[lane] = L.y;
    st->save[lane] = L.save;
#pragma unroll
    for (int k = 0; k < O; ++k) st->hist[k][lane] = L.hist[k];
    if (lane == 0) {
        st->t = L.t;
        st->dt = L.dt;
        st->yield_memory = L.yield_memory;
        st->hist_full = L.hist_full ? 1 : 0;
#pragma unroll
        for (int k = 0; k < O; ++k) st->hist_t[k] = L.hist_t[k];
        st->steps = steps;
        st->n_accepted = nacc;
        st->n_rejected = nrej;
        st->n_newton = L.n_newton;
        st->n_lu = L.n_lu;
        st->status = status;
        st->finished = finished;
        st->count = count;
    }
}

template <int = 0>
__global__ void bdf_setup_kernel(uint32_t *counter, bg_ensemble_desc *dst, const bg_ensemble_desc d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *dst = d;
    }
}

// ---------------------------------------------------------------------------
// The corrector's inverse cascade on caller matrices (bg_bdf_inverse_device):
// one matrix per group, W x W row-major in `a`; lane (h, g) stages row g in
// the group's LDS rows, runs partial-pivot Gauss-Jordan, and on failure the
// full-pivot and QR members on the reloaded matrix, exactly as the secant
// does (from member `first` on: 0 the whole cascade, 1 full pivot, 2 QR);
// it writes its part of row g of the inverse and the member that answered
// (0, 1, 2; -1: SingularMatrix).
// ---------------------------------------------------------------------------
template <int W_>
struct InverseOnly {  // a right-hand side in name only: the lane layout of W_
    static constexpr int D = W_, W = W_, P = 0;
};

template <int W, int SPLIT>
__global__ void __launch_bounds__(64)
    bdf_inverse_kernel(const double *__restrict__ a, double *__restrict__ inv,
                       int32_t *__restrict__ branch, uint64_t count, int first) {
    constexpr int GW = W * SPLIT, GPW = 64 / GW;
    __shared__ double lds_all[GPW][W * (W + 1) + 2];
    const int lane = __lane_id();
    const int grp_id = lane / GW;
    BdfLane<InverseOnly<W>, 7, BdfSem<false>, double, SPLIT> L;
    L.grp.g = lane % W;
    L.h = (lane % GW) / W;
    L.lds = lds_all[grp_id];
    const uint64_t m = (uint64_t)blockIdx.x * GPW + grp_id;
    const bool live = m < count;
    const uint64_t mm = live ? m : 0;  // a spare group inverts matrix 0 and writes nothing
    auto stage = [&]() {
        for (int c = 0; c < W; ++c)
            L.lds[L.grp.g * (W + 1) + c] = a[(mm * W + L.grp.g) * W + c];
        __builtin_amdgcn_wave_barrier();
    };
    stage();
    int br = first;
    bool ok = false;
    if (first == 0) ok = L.invert();
    if (!ok && first <= 1) {
        stage();
        br = 1;
        ok = L.invert_full_pivot();
    }
    if (!ok) {
        stage();
        br = 2;
        ok = L.invert_qr();
    }
    if (!ok) br = -1;
    if (!live) return;
#pragma unroll
    for (int q = 0; q < L.WJ; ++q) inv[(mm * W + L.grp.g) * W + L.h * L.WJ + q] = L.jinv[q];
    if (L.grp.g == 0 && L.h == 0) branch[mm] = br;
}

// ---------------------------------------------------------------------------
// launch (each translation unit instantiates its own set of systems)
// ---------------------------------------------------------------------------
template <class T>
struct BTag {
    using type = T;
};

// Persistent grid (one wave per block, as many as are resident, at most
// cap_per_cu per CU when cap_per_cu > 0) and the launch workspace: the work
// counter and the descriptor copy.  *occupancy, if given, receives the
// resident blocks per CU before the cap.
template <class R, int O, class S, class F>
static int bdf_launch_shape(const bg_ensemble_desc &d, uint64_t &grid, size_t &bytes,
                            uint32_t cap_per_cu = 0, int *occupancy = nullptr) {
    constexpr int W = R::W;
    constexpr int SP = W == 32 ? 2 : 1;  // a 32-component instance takes the whole wave
    constexpr int GPW = 64 / (W * SP);
    int dev = 0;
    BG_HIP_CHECK(hipGetDevice(&dev));
    int per_cu = 0;
    BG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, bdf_ensemble_kernel<R, O, S, F, SP>, 64, 0));
    if (occupancy) *occupancy = per_cu;
    if (cap_per_cu && (int)cap_per_cu < per_cu) per_cu = (int)cap_per_cu;
    if (per_cu < 1) per_cu = 1;
    grid = (uint64_t)per_cu * (uint64_t)device_cu_count(dev);
    const uint64_t need = (d.n + GPW - 1) / GPW;
    if (grid > need) grid = need;
    bytes = kLaunchWorkspaceBytes;
    return BG_OK;
}

template <class R, int O, class S, class F>
static int bdf_launch_one(const bg_ensemble_desc &d, hipStream_t stream, void *ws_in,
                          size_t ws_bytes, uint32_t cap_per_cu = 0) {
    if (d.n == 0) return BG_OK;
    if (d.n > 0x7fffffffull || !d.y0) return BG_ERR_INVALID_ARGUMENT;
    if (R::P > 0 && (!d.params || d.n_params < (uint32_t)R::P)) return BG_ERR_INVALID_ARGUMENT;
    constexpr int SP = R::W == 32 ? 2 : 1;
    uint64_t grid = 0;
    size_t bytes = 0;
    const int rc = bdf_launch_shape<R, O, S, F>(d, grid, bytes, cap_per_cu);
    if (rc != BG_OK) return rc;
    if (ws_in && ws_bytes < bytes) return BG_ERR_INVALID_ARGUMENT;
    char *ws = (char *)ws_in;
    if (!ws_in) BG_HIP_CHECK(hipMallocAsync((void **)&ws, bytes, stream));
    uint32_t *counter = (uint32_t *)ws;
    bg_ensemble_desc *ddesc = (bg_ensemble_desc *)(ws + 256);
    bdf_setup_kernel<<<1, 64, 0, stream>>>(counter, ddesc, d);
    bdf_ensemble_kernel<R, O, S, F, SP><<<(uint32_t)grid, 64, 0, stream>>>(ddesc, counter);
    const hipError_t launched = hipGetLastError();
    if (!ws_in) BG_HIP_CHECK(hipFreeAsync(ws, stream));
    BG_HIP_CHECK(launched);
    return BG_OK;
}

template <class R, int O, class S, class F>
static int bdf_path_one(BdfPathState *state, double *path_t, double *path_y, uint32_t cap,
                        hipStream_t stream) {
    bdf_path_kernel<R, O, S, F><<<1, 64, 0, stream>>>(state, path_t, path_y, cap);
    BG_HIP_CHECK(hipGetLastError());
    return BG_OK;
}

// Method and semantics dispatch around a per-system visitor `sys(rtag,
// ftag)` of the caller's translation unit.
template <class V>
static int bdf_with_order_sem(int method, int semantics, V &&v) {
    if (semantics != BG_AS_BUILT && semantics != BG_INTENDED) return BG_ERR_INVALID_ARGUMENT;
    auto with_sem = [&](auto otag) -> int {
        return semantics == BG_AS_BUILT ? v(otag, BTag<BdfSem<true>>{})
                                        : v(otag, BTag<BdfSem<false>>{});
    };
    if (method == BG_BDF6) return with_sem(std::integral_constant<int, 7>{});
    if (method == BG_BDF2) return with_sem(std::integral_constant<int, 3>{});
    return BG_ERR_UNSUPPORTED;
}

}  // namespace bg
