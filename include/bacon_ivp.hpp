// bacon_ivp.hpp — C++17 host-side mirror of bacon-sci's `ivp` builder API over
// the C ABI in bacon_gpu.h (header only; link with libbacon_gpu.so).
//
// Names, argument meaning and error behaviour follow the reference:
//   IVPError                    src/ivp.rs:50-76 (thrown as bacon::ivp::IVPError)
//   IVPSolver setters           src/ivp/rk.rs:168-247 (bdf.rs:176-255 identical)
//   solve(data)                 src/ivp/rk.rs:249-343
//   IVPIterator / collect_vec   src/ivp.rs:192-239
//   RungeKutta45 / RungeKutta23 src/ivp/rk.rs:561, 656
//   BDF6 / BDF2                 src/ivp/bdf.rs:706, 762
//   Adams5 / Adams3             src/ivp/adams.rs:633, 693
// A Rust setter returns Result<Self, IVPError>; here it returns *this or
// throws IVPError, so builder chains read the same.  The derivative is a
// registered device function id plus its parameters.
//
//   auto path = bacon::ivp::RungeKutta45::make(1)
//                   .with_minimum_dt(1e-4).with_maximum_dt(0.1)
//                   .with_initial_time(0.0).with_ending_time(10.0)
//                   .with_tolerance(1e-5)
//                   .with_initial_conditions_slice({1.0})
//                   .with_derivative(BG_RHS_QUADRATIC)
//                   .solve()
//                   .collect_vec();
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bacon_gpu.h"

namespace bacon {
namespace ivp {

inline const char *error_name(int code) {
    switch (code) {
    case BG_MISSING_PARAMETERS: return "MissingParameters";
    case BG_USER_ERROR: return "UserError";
    case BG_TOLERANCE_OOB: return "ToleranceOOB";
    case BG_TIME_DELTA_OOB: return "TimeDeltaOOB";
    case BG_TIME_END_OOB: return "TimeEndOOB";
    case BG_TIME_START_OOB: return "TimeStartOOB";
    case BG_FROM_PRIMITIVE_FAILURE: return "FromPrimitiveFailure";
    case BG_MINIMUM_TIME_DELTA_EXCEEDED: return "MinimumTimeDeltaExceeded";
    case BG_MAXIMUM_ITERATIONS_EXCEEDED: return "MaximumIterationsExceeded";
    case BG_SINGULAR_MATRIX: return "SingularMatrix";
    case BG_DYNAMIC_ON_STATIC: return "DynamicOnStatic";
    case BG_STATIC_ON_DYNAMIC: return "StaticOnDynamic";
    case BG_ERR_INVALID_ARGUMENT: return "InvalidArgument";
    case BG_ERR_HIP: return "HipError";
    case BG_ERR_UNSUPPORTED: return "Unsupported";
    case BG_ERR_STATE: return "StateError";
    default: return "UnknownError";
    }
}

// IVPError (src/ivp.rs:50-76); code() is the C ABI status.
class IVPError : public std::runtime_error {
  public:
    explicit IVPError(int code) : std::runtime_error(error_name(code)), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc != BG_OK) throw IVPError(rc);
}

using Point = std::pair<double, std::vector<double>>;

// IVPIterator (src/ivp.rs:192-239): next() yields (t, y), returns nullopt at
// the end, throws the solver's failure once, then stays exhausted.
class IVPIterator {
  public:
    explicit IVPIterator(std::shared_ptr<bg_solver> s, uint32_t dim) : s_(std::move(s)), dim_(dim) {}

    std::optional<Point> next() {
        double t = 0.0;
        std::vector<double> y(dim_);
        const int rc = bg_next(s_.get(), &t, y.data());
        if (rc == BG_OK) return Point{t, std::move(y)};
        if (rc == BG_DONE) return std::nullopt;
        throw IVPError(rc);
    }

    // Path::collect_vec (src/ivp.rs:209-211)
    std::vector<Point> collect_vec() {
        std::vector<Point> out;
        while (auto p = next()) out.push_back(std::move(*p));
        return out;
    }

    double time() const {
        double t = 0.0;
        check(bg_time(s_.get(), &t));
        return t;
    }

  private:
    std::shared_ptr<bg_solver> s_;
    uint32_t dim_;
};

template <int METHOD>
class Solver {
  public:
    // new() for a solver whose dimension is static (rk.rs:136-148)
    static Solver make(uint32_t dim, int dim_kind = BG_DIM_STATIC) {
        bg_solver *h = nullptr;
        check(bg_new(METHOD, dim_kind, dim, &h));
        return Solver(h);
    }
    // new_dyn(size) (rk.rs:150-162)
    static Solver make_dyn(uint32_t size, int dim_kind = BG_DIM_DYN) {
        bg_solver *h = nullptr;
        check(bg_new_dyn(METHOD, dim_kind, size, &h));
        return Solver(h);
    }

    uint32_t dim() const {
        uint32_t d = 0;
        check(bg_dim(h_.get(), &d));
        return d;
    }
    Solver &with_tolerance(double tol) { return set(bg_with_tolerance(h_.get(), tol)); }
    Solver &with_maximum_dt(double v) { return set(bg_with_maximum_dt(h_.get(), v)); }
    Solver &with_minimum_dt(double v) { return set(bg_with_minimum_dt(h_.get(), v)); }
    Solver &with_initial_time(double v) { return set(bg_with_initial_time(h_.get(), v)); }
    Solver &with_ending_time(double v) { return set(bg_with_ending_time(h_.get(), v)); }
    Solver &with_initial_conditions_slice(const std::vector<double> &y0) {
        return set(bg_with_initial_conditions_slice(h_.get(), y0.data(), y0.size()));
    }
    Solver &with_derivative(int rhs, const std::vector<double> &params = {}) {
        return set(bg_with_derivative(h_.get(), rhs, params.empty() ? nullptr : params.data(),
                                      params.size()));
    }
    Solver &with_semantics(int semantics) { return set(bg_with_semantics(h_.get(), semantics)); }
    // README.md:33-38 spellings
    Solver &with_dt_min(double v) { return set(bg_with_dt_min(h_.get(), v)); }
    Solver &with_dt_max(double v) { return set(bg_with_dt_max(h_.get(), v)); }
    Solver &with_start(double v) { return set(bg_with_start(h_.get(), v)); }
    Solver &with_end(double v) { return set(bg_with_end(h_.get(), v)); }

    // solve(data) (rk.rs:249-343)
    IVPIterator solve() {
        check(bg_solve(h_.get()));
        return IVPIterator(h_, dim());
    }

  private:
    explicit Solver(bg_solver *h) : h_(h, bg_destroy) {}
    Solver &set(int rc) {
        check(rc);
        return *this;
    }
    std::shared_ptr<bg_solver> h_;
};

using RungeKutta45 = Solver<BG_RK45>;
using RungeKutta23 = Solver<BG_RK23>;
using RK45 = RungeKutta45;  // README.md:23-47 spelling
using BDF6 = Solver<BG_BDF6>;
using BDF2 = Solver<BG_BDF2>;
using Adams5 = Solver<BG_ADAMS5>;
using Adams3 = Solver<BG_ADAMS3>;

}  // namespace ivp
}  // namespace bacon
