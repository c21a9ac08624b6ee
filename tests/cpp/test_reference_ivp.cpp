// The reference's own ivp tests (src/ivp/rk.rs:682-758, src/ivp/bdf.rs:785-923)
// restated against the C++ mirror of the builder API (include/bacon_ivp.hpp),
// run on the GPU through libbacon_gpu.so.  Exit status 0 = all passed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>

#include "../../include/bacon_ivp.hpp"

using namespace bacon::ivp;

static int failures = 0;
#define EXPECT(cond, what)                                      \
    do {                                                        \
        if (!(cond)) {                                          \
            std::printf("FAIL %s: %s\n", what, #cond);          \
            ++failures;                                         \
        }                                                       \
    } while (0)

// float_cmp::approx_eq!(f64, a, b, epsilon = eps): |a-b| <= eps or <= 4 ULPs
static bool approx_eq(double a, double b, double eps) {
    if (std::fabs(a - b) <= eps) return true;
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    if ((ia < 0) != (ib < 0)) return false;
    const int64_t d = ia > ib ? ia - ib : ib - ia;
    return d <= 4;
}

template <class S>
static void rungekutta_quadratic(const char *name, int semantics) {
    // rk.rs:682-719
    auto path = S::make(1)
                    .with_minimum_dt(0.0001)
                    .with_maximum_dt(0.1)
                    .with_initial_time(0.0)
                    .with_ending_time(10.0)
                    .with_tolerance(1e-5)
                    .with_initial_conditions_slice({1.0})
                    .with_derivative(BG_RHS_QUADRATIC)
                    .with_semantics(semantics)
                    .solve()
                    .collect_vec();
    EXPECT(!path.empty(), name);
    for (auto &p : path) EXPECT(approx_eq(p.second[0], 1.0 - p.first * p.first, 0.0001), name);
}

template <class S>
static void rungekutta_sine(const char *name, int semantics) {
    // rk.rs:721-758
    auto path = S::make(1)
                    .with_minimum_dt(0.001)
                    .with_maximum_dt(0.01)
                    .with_tolerance(0.0001)
                    .with_initial_time(0.0)
                    .with_ending_time(10.0)
                    .with_initial_conditions_slice({0.0})
                    .with_derivative(BG_RHS_COS)
                    .with_semantics(semantics)
                    .solve()
                    .collect_vec();
    EXPECT(!path.empty(), name);
    for (auto &p : path) EXPECT(approx_eq(p.second[0], std::sin(p.first), 0.01), name);
}

template <class S>
static void bdf_case(const char *name, int rhs, double y0, double t_end,
                     std::function<double(double)> exact, int semantics, bool expect_empty) {
    // bdf.rs:785-923
    auto path = S::make(1)
                    .with_minimum_dt(1e-5)
                    .with_maximum_dt(0.1)
                    .with_tolerance(0.00001)
                    .with_initial_time(0.0)
                    .with_ending_time(t_end)
                    .with_initial_conditions_slice({y0})
                    .with_derivative(rhs)
                    .with_semantics(semantics)
                    .solve()
                    .collect_vec();
    if (expect_empty) EXPECT(path.empty(), name);
    else EXPECT(!path.empty(), name);
    for (auto &p : path) EXPECT(approx_eq(p.second[0], exact(p.first), 0.01), name);
}

template <class S>
static void adams_case(const char *name, int rhs, double y0, double t_end, double dt_min,
                       double dt_max, double tol, std::function<double(double)> exact,
                       int semantics) {
    // adams.rs:713-922
    auto path = S::make(1)
                    .with_minimum_dt(dt_min)
                    .with_maximum_dt(dt_max)
                    .with_tolerance(tol)
                    .with_initial_time(0.0)
                    .with_ending_time(t_end)
                    .with_initial_conditions_slice({y0})
                    .with_derivative(rhs)
                    .with_semantics(semantics)
                    .solve()
                    .collect_vec();
    EXPECT(!path.empty(), name);
    for (auto &p : path) EXPECT(approx_eq(p.second[0], exact(p.first), 0.01), name);
}

static void builder_errors() {
    try {
        RK45::make(1).with_tolerance(0.0);
        EXPECT(false, "ToleranceOOB not thrown");
    } catch (const IVPError &e) {
        EXPECT(e.code() == BG_TOLERANCE_OOB, "ToleranceOOB");
    }
    try {
        RK45::make(1).with_ending_time(1.0).with_initial_time(2.0);
        EXPECT(false, "TimeStartOOB not thrown");
    } catch (const IVPError &e) {
        EXPECT(e.code() == BG_TIME_START_OOB, "TimeStartOOB");
    }
    try {
        RK45::make(1).with_tolerance(1e-3).solve();
        EXPECT(false, "MissingParameters not thrown");
    } catch (const IVPError &e) {
        EXPECT(e.code() == BG_MISSING_PARAMETERS, "MissingParameters");
    }
    try {
        RK45::make(1, BG_DIM_DYN);
        EXPECT(false, "StaticOnDynamic not thrown");
    } catch (const IVPError &e) {
        EXPECT(e.code() == BG_STATIC_ON_DYNAMIC, "StaticOnDynamic");
    }
    // the iterator reports a failure once, then ends (ivp.rs:232-235)
    auto it = RK45::make(1)
                  .with_minimum_dt(1e-6)
                  .with_maximum_dt(0.01)
                  .with_tolerance(1e-8)
                  .with_initial_time(0.0)
                  .with_ending_time(2.0)
                  .with_initial_conditions_slice({1.0})
                  .with_derivative(BG_RHS_FAIL_AFTER, {0.5})
                  .solve();
    int yielded = 0, errors = 0;
    for (;;) {
        try {
            auto p = it.next();
            if (!p) break;
            ++yielded;
        } catch (const IVPError &e) {
            EXPECT(e.code() == BG_USER_ERROR, "UserError");
            ++errors;
        }
    }
    EXPECT(yielded > 0 && errors == 1, "fused iterator");
}

int main() {
    for (int sem : {BG_AS_BUILT, BG_INTENDED}) {
        rungekutta_quadratic<RungeKutta45>("rk45 quadratic", sem);
        rungekutta_quadratic<RungeKutta23>("rk23 quadratic", sem);
        rungekutta_sine<RungeKutta45>("rk45 sine", sem);
        rungekutta_sine<RungeKutta23>("rk23 sine", sem);
        const bool empty = sem == BG_AS_BUILT;  // bdf.rs:622 ends the solve
        bdf_case<BDF6>("bdf6 exp", BG_RHS_EXP, 1.0, 7.0, [](double t) { return std::exp(t); }, sem, empty);
        bdf_case<BDF6>("bdf6 unstable", BG_RHS_NEG, 1.0, 10.0, [](double t) { return std::exp(-t); }, sem, empty);
        bdf_case<BDF6>("bdf6 quadratic", BG_RHS_QUADRATIC, 1.0, 2.0, [](double t) { return 1.0 - t * t; }, sem, empty);
        bdf_case<BDF6>("bdf6 sin", BG_RHS_COS, 0.0, 6.0, [](double t) { return std::sin(t); }, sem, empty);
        bdf_case<BDF2>("bdf2 exp", BG_RHS_EXP, 1.0, 7.0, [](double t) { return std::exp(t); }, sem, empty);
        bdf_case<BDF2>("bdf2 quadratic", BG_RHS_QUADRATIC, 1.0, 1.0, [](double t) { return 1.0 - t * t; }, sem, empty);
        const double tau = 6.283185307179586;
        auto ex = [](double t) { return std::exp(t); };
        auto quad = [](double t) { return 1.0 - t * t; };
        auto sine = [](double t) { return std::sin(t); };
        adams_case<Adams5>("adams5 exp", BG_RHS_EXP, 1.0, 2.0, 1e-5, 0.1, 0.0005, ex, sem);
        adams_case<Adams5>("adams5 quadratic", BG_RHS_QUADRATIC, 1.0, 5.0, 1e-7, 0.001, 0.01, quad, sem);
        adams_case<Adams5>("adams5 sine", BG_RHS_COS, 0.0, tau, 1e-5, 0.001, 0.01, sine, sem);
        adams_case<Adams3>("adams3 exp", BG_RHS_EXP, 1.0, 2.0, 1e-5, 0.1, 0.001, ex, sem);
        adams_case<Adams3>("adams3 quadratic", BG_RHS_QUADRATIC, 1.0, 5.0, 1e-5, 0.001, 0.1, quad, sem);
        adams_case<Adams3>("adams3 sine", BG_RHS_COS, 0.0, tau, 1e-5, 0.001, 0.01, sine, sem);
    }
    builder_errors();
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
