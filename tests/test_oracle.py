"""Pins the CPU oracle (oracle/bacon_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake (a wrong coefficient, a
transposed matrix, a dropped term, a wrong sign or index) fails it:
  * the reference's own ivp tests (src/ivp/rk.rs:682-758, src/ivp/bdf.rs:785-1063)
    with float_cmp's approx_eq semantics;
  * classical order conditions of the Fehlberg pair and the BDF formulas;
  * closed forms (polynomial exactness, the linear-test-function amplification
    R(z) in exact rationals, the as-built first step = explicit Euler, the
    controller's dt sequence on an exactly integrated RHS);
  * independent high-accuracy scipy solutions (tests/golden/*.npz), an exact
    matrix-exponential solution for the linear workload, and the conserved
    quantity of Lotka-Volterra;
  * LAPACK inverses for each member of the nalgebra inverse cascade (LU,
    full-pivot LU, QR) and constructed inputs that fix which member answers.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from bacon_amd import workloads as W


def approx_eq(a, b, eps):
    """float_cmp::approx_eq!(f64, a, b, epsilon = eps): |a-b| <= eps or <= 4 ULPs."""
    if abs(a - b) <= eps:
        return True
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    return (ia < 0) == (ib < 0) and abs(int(ia) - int(ib)) <= 4


# ---------------------------------------------------------------- reference tests
@pytest.mark.parametrize("semantics", [O.AS_BUILT, O.INTENDED])
@pytest.mark.parametrize("method", [O.RK45, O.RK23])
def test_reference_rungekutta_quadratic(orc, method, semantics):
    # rk.rs:682-719
    r = O.solve(method, "quadratic", [1.0], 0.0, 10.0, 1e-5, 1e-4, 0.1,
                semantics=semantics, path_cap=100000)
    assert r["status"] == O.OK
    assert r["path_len"] > 0
    for t, y in zip(r["path_t"], r["path_y"]):
        assert approx_eq(y[0], 1.0 - t * t, 1e-4)


@pytest.mark.parametrize("semantics", [O.AS_BUILT, O.INTENDED])
@pytest.mark.parametrize("method", [O.RK45, O.RK23])
def test_reference_rungekutta_sine(orc, method, semantics):
    # rk.rs:721-758
    r = O.solve(method, "cos", [0.0], 0.0, 10.0, 1e-4, 1e-3, 1e-2,
                semantics=semantics, path_cap=100000)
    assert r["status"] == O.OK
    assert r["path_len"] > 0
    for t, y in zip(r["path_t"], r["path_y"]):
        assert approx_eq(y[0], np.sin(t), 0.01)


BDF_CASES = [
    # name, rhs, y0, t_end (BDF6, BDF2), exact       (bdf.rs:785-1063)
    ("exp", "exp", 1.0, (7.0, 7.0), np.exp),
    ("unstable", "neg", 1.0, (10.0, 10.0), lambda t: np.exp(-t)),
    ("quadratic", "quadratic", 1.0, (2.0, 1.0), lambda t: 1.0 - t * t),
    ("sin", "cos", 0.0, (6.0, 6.0), np.sin),
]


@pytest.mark.parametrize("case", BDF_CASES, ids=[c[0] for c in BDF_CASES])
@pytest.mark.parametrize("method", [O.BDF6, O.BDF2])
def test_reference_bdf_as_built_yields_nothing(orc, method, case):
    """As built, the first BDF step after the RK4 bootstrap is rejected and the
    restore `time -= dt - order` (bdf.rs:622) moves the clock past t_end, so
    the reference's own BDF tests iterate over an empty path and pass
    vacuously."""
    name, rhs, y0, tends, exact = case
    t_end = tends[0] if method == O.BDF6 else tends[1]
    r = O.solve(method, rhs, [y0], 0.0, t_end, 1e-5, 1e-5, 0.1,
                semantics=O.AS_BUILT, path_cap=1000)
    assert r["status"] == O.OK
    assert r["path_len"] == 0
    assert r["t"] > t_end


@pytest.mark.parametrize("case", BDF_CASES, ids=[c[0] for c in BDF_CASES])
@pytest.mark.parametrize("method", [O.BDF6, O.BDF2])
def test_reference_bdf_intended(orc, method, case):
    name, rhs, y0, tends, exact = case
    t_end = tends[0] if method == O.BDF6 else tends[1]
    r = O.solve(method, rhs, [y0], 0.0, t_end, 1e-5, 1e-5, 0.1,
                semantics=O.INTENDED, path_cap=200000)
    assert r["status"] == O.OK
    assert r["path_len"] > 0
    # the solver reaches t_end; the last yield may be earlier, because RK4
    # bootstrap values are not yielded when the bootstrap ends the solve
    # (bdf.rs:531-546)
    assert abs(r["t"] - t_end) < 1e-9
    assert approx_eq(r["y"][0], exact(r["t"]), 0.01)
    for t, y in zip(r["path_t"], r["path_y"]):
        assert approx_eq(y[0], exact(t), 0.01)


# ---------------------------------------------------------------- tableau
def _elementary_weights(c, a, b):
    """The eight order-4 conditions for an explicit RK method."""
    return np.array([
        b.sum() - 1,
        b @ c - 1 / 2,
        b @ c**2 - 1 / 3,
        b @ a @ c - 1 / 6,
        b @ c**3 - 1 / 4,
        (b * c) @ a @ c - 1 / 8,
        b @ a @ c**2 - 1 / 12,
        b @ a @ a @ c - 1 / 24,
    ])


def _fixed_step(c, a, b, h, t_end=2.0):
    """Pendulum y0' = y1, y1' = -sin(y0), y(0) = (1, 0), fixed steps of size h."""
    f = lambda y: np.array([y[1], -np.sin(y[0])])  # noqa: E731
    y = np.array([1.0, 0.0])
    s = len(c)
    for _ in range(int(round(t_end / h))):
        k = np.zeros((s, 2))
        for i in range(s):
            k[i] = f(y + h * (a[i, :i] @ k[:i]))
        y = y + h * (b @ k)
    return y


def test_fehlberg_intended_order_conditions(orc):
    c, m, b, e = O.rk_tableau(O.RK45, O.INTENDED)
    assert np.allclose(np.tril(m, -1), m)  # explicit
    assert np.allclose(m.sum(axis=1), c, atol=1e-15)  # row-sum condition
    assert np.abs(_elementary_weights(c, m, b)).max() < 1e-15
    assert abs(e.sum()) < 1e-16
    b5 = b + e
    assert np.abs(_elementary_weights(c, m, b5)).max() < 1e-15
    # observed convergence orders: 4 for the propagated weights, 5 for b + e
    ref = _fixed_step(c, m, b5, 0.0005)
    for w, p in ((b, 4), (b5, 5)):
        e1 = np.abs(_fixed_step(c, m, w, 0.05) - ref).max()
        e2 = np.abs(_fixed_step(c, m, w, 0.025) - ref).max()
        assert abs(np.log2(e1 / e2) - p) < 0.2, (p, np.log2(e1 / e2))


def test_bogacki_shampine_order(orc):
    c, m, b, e = O.rk_tableau(O.RK23, O.INTENDED)
    assert np.allclose(m.sum(axis=1), c)
    assert abs(b.sum() - 1) < 1e-15 and abs(b @ c - 0.5) < 1e-15
    assert abs(b @ c**2 - 1 / 3) < 1e-15 and abs(b @ m @ c - 1 / 6) < 1e-15
    assert abs(e.sum()) < 1e-16


def test_as_built_reads_transpose(orc):
    _, mi, _, _ = O.rk_tableau(O.RK45, O.INTENDED)
    _, ma, _, _ = O.rk_tableau(O.RK45, O.AS_BUILT)
    assert np.array_equal(ma, mi.T)
    assert np.array_equal(np.triu(ma, 1), ma)  # stages read later (stale) stages


def _R_exact(c, a, b, z):
    """Stability function R(z) = 1 + z b^T (I - zA)^{-1} 1 in exact rationals."""
    s = len(c)
    A = [[Fraction(a[i][j]) for j in range(s)] for i in range(s)]
    B = [Fraction(x) for x in b]
    k = [Fraction(0)] * s
    for i in range(s):
        k[i] = z * (1 + sum(A[i][j] * k[j] for j in range(i)))
    return 1 + sum(B[i] * k[i] for i in range(s))


def test_linear_amplification_intended(orc):
    # one accepted step of y' = -y with dt0 = h (dt_min = dt_max = h)
    h = 0.0625
    r = O.solve(O.RK45, "neg", [1.0], 0.0, 1.0, 1e-3, h, h, semantics=O.INTENDED,
                path_cap=4)
    c, m, b, e = O.rk_tableau(O.RK45, O.INTENDED)
    R = _R_exact(c, m, b, Fraction(-h))
    assert r["path_dt"][0] == h
    assert abs(r["path_y"][0][0] - float(R)) < 4e-16


def test_as_built_first_step_is_euler(orc):
    """With all stages zero on entry, every as-built stage is evaluated at y."""
    dt0 = (0.1 + 1e-4) / 2
    r = O.solve(O.RK45, "exp", [1.0], 0.0, 1.0, 1e-3, 1e-4, 0.1, semantics=O.AS_BUILT,
                path_cap=4)
    _, _, b, _ = O.rk_tableau(O.RK45, O.AS_BUILT)
    assert r["path_t"][0] == dt0
    assert abs(r["path_y"][0][0] - (1.0 + dt0 * b.sum())) < 4e-16
    # Euler's local error is invisible to the estimate (sum(e) ~ 0), so the
    # controller grows dt by the maximum factor 4, capped at dt_max
    r1 = O.solve(O.RK45, "exp", [1.0], 0.0, 1.0, 1e-3, 1e-4, 0.1, semantics=O.AS_BUILT,
                 max_attempts=1)
    assert r1["n_accepted"] == 1 and r1["dt"] == min(4 * dt0, 0.1)


@pytest.mark.parametrize("semantics", [O.AS_BUILT, O.INTENDED])
def test_controller_dt_sequence_on_exact_rhs(orc, semantics):
    """y' = -2t is integrated exactly, the error is zero, delta = inf, so dt
    grows by 4 and is capped at dt_max (rk.rs:400-412); the last step is
    clamped to the end (rk.rs:366-368)."""
    r = O.solve(O.RK45, "quadratic", [1.0], 0.0, 1.0, 1e-6, 1e-4, 0.1,
                semantics=semantics, path_cap=100)
    dt0 = (0.1 + 1e-4) / 2
    t = 0.0
    expect_t, expect_dt = [], []
    dt = dt0
    while t < 1.0:
        if t + dt >= 1.0:
            dt = 1.0 - t
        t += dt
        expect_t.append(t)
        expect_dt.append(dt)
        dt = min(dt * 4.0, 0.1)
    assert np.array_equal(r["path_t"], np.array(expect_t))
    assert np.array_equal(r["path_dt"], np.array(expect_dt))
    assert np.abs(r["path_y"][:, 0] - (1 - r["path_t"] ** 2)).max() < 1e-14


def test_as_built_lorenz_fails_min_dt(orc):
    """Q1 analysis: for a y-dependent RHS the as-built method is first order and
    its estimate ~0.277 h |J f| forces dt below dt_min on the plain Lorenz
    workload: MinimumTimeDeltaExceeded on every instance after one step."""
    w = W.lorenz_plain(64)
    r = O.solve_ensemble(O.RK45, "lorenz", w["y0"], w["params"], w["t_end"], w["tol"],
                         dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.AS_BUILT)
    assert np.all(r["status"] == O.MINIMUM_TIME_DELTA_EXCEEDED)
    assert np.all(r["n_accepted"] == 1)


def test_user_error_and_fuse(orc):
    r = O.solve(O.RK45, "fail_after", [1.0], 0.0, 2.0, 1e-8, 1e-6, 0.01, params=[0.5],
                path_cap=10000)
    assert r["status"] == O.USER_ERROR
    assert r["path_len"] > 0 and r["path_t"][-1] <= 0.5 + 0.01


def test_nan_error_hits_attempt_cap(orc):
    """A NaN error makes the reference redo forever (rk.rs:392-422); the cap
    reports MaximumIterationsExceeded."""
    r = O.solve(O.RK45, "lorenz", [1e200, -1e200, 1e200], 0.0, 1.0, 1e-6, 1e-6, 0.1,
                params=[10.0, 28.0, 8.0 / 3.0], max_attempts=500)
    assert r["status"] == O.MAXIMUM_ITERATIONS_EXCEEDED


# ---------------------------------------------------------------- BDF coefficients
@pytest.mark.parametrize("method,order", [(O.BDF6, 6), (O.BDF2, 2)])
def test_bdf_order_conditions(orc, method, order):
    hi, lo = O.bdf_coefficients(method)
    alpha = np.concatenate([[1.0], hi[1:]])  # node -k for k = 0..O-1
    nodes = -np.arange(alpha.size, dtype=float)
    assert abs(alpha.sum()) < 1e-14
    assert abs((alpha * nodes).sum() - hi[0]) < 1e-14
    for q in range(2, order + 1):
        assert abs((alpha * nodes**q).sum()) < 1e-11, q
    # the lower formula is one order lower and consistent with its own beta0
    alpha_l = np.concatenate([[1.0], lo[1:]])
    nodes_l = -np.arange(alpha_l.size, dtype=float)
    assert abs(alpha_l.sum()) < 1e-14
    assert abs((alpha_l * nodes_l).sum() - lo[0]) < 1e-14
    for q in range(2, order):
        assert abs((alpha_l * nodes_l**q).sum()) < 1e-11, q


# ---------------------------------------------------------------- nalgebra LU
def test_lu_inverse_matches_lapack(orc):
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 8, 32):
        a = rng.standard_normal((n, n))
        rc, inv, which = O.inverse(a)
        assert rc == O.OK and which == 0
        ref = np.linalg.inv(a)
        assert np.abs(inv - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


def test_inverse_singular(orc):
    a = np.ones((4, 4))
    rc, _, which = O.inverse(a)
    assert rc == O.SINGULAR_MATRIX and which == -1
    a = np.zeros((3, 3))
    rc, _, _ = O.inverse(a)
    assert rc == O.SINGULAR_MATRIX


def test_inverse_needs_pivoting(orc):
    a = np.array([[0.0, 1.0], [1.0, 0.0]])
    rc, inv, which = O.inverse(a)
    assert rc == O.OK and np.array_equal(inv, a)


def _pivot_heavy(rng, n):
    """A random matrix whose largest entries sit off the diagonal in shuffled
    rows and columns, so row and column pivots both move at every stage."""
    a = rng.standard_normal((n, n))
    a[rng.permutation(n), rng.permutation(n)] += 8.0 * np.sign(rng.standard_normal(n))
    return a


@pytest.mark.parametrize("branch", [O.INVERSE_LU, O.INVERSE_FULL_PIVOT, O.INVERSE_QR])
def test_each_inverse_matches_lapack(orc, branch):
    """Every member of the cascade (bdf.rs:429-441: lu, full_piv_lu, qr) on
    its own against LAPACK, at every size 1..32, on plain Gaussian matrices
    and on matrices that force row and column interchanges.  A wrong sign in
    a reflection, a permutation applied in the wrong order or to the wrong
    side, or a dropped update fails this."""
    rng = np.random.default_rng(11 + branch)
    for n in range(1, 33):
        for a in (rng.standard_normal((n, n)), _pivot_heavy(rng, n)):
            rc, inv = O.inverse_branch(a, branch)
            assert rc == O.OK, (branch, n)
            ref = np.linalg.inv(a)
            scale = np.linalg.cond(a) * max(1.0, np.abs(ref).max())
            assert np.abs(inv - ref).max() <= 1e-13 * n * scale, (branch, n)


def _residual_ratio(a, x):
    """max |A X - I| / (|A| |X|): small for a backward-stable inverse whatever
    the conditioning."""
    return np.abs(a @ x - np.eye(a.shape[0])).max() / (
        np.abs(a).sum(1).max() * np.abs(x).sum(0).max())


def test_inverse_cascade_branches(orc):
    """Which member of the cascade answers (bdf.rs:429-444).  The inputs are
    built so that plain double arithmetic, restated here independently of the
    oracle, fixes the outcome of each factorisation."""
    # 1. well conditioned: partial-pivot LU answers
    a = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.25], [0.5, 0.25, 2.0]])
    rc, inv, which = O.inverse(a)
    assert rc == O.OK and which == 0
    assert np.abs(inv - np.linalg.inv(a)).max() < 1e-14

    # 2. partial-pivot LU meets an exact zero pivot, full-pivot LU does not.
    #    [[p, b], [c, d]] with |p| > |c|: LU pivots on p, the multiplier is
    #    l = c * (1/p), and the trailing entry -b*l + d is exactly 0 when
    #    d = fl(b*l).  Full pivoting takes b (the largest entry) first; after
    #    the column interchange the trailing entry is -p*l2 + c with
    #    l2 = d * (1/b), which is not zero.
    p, b, c = 26.0, 48.0, 23.0
    l = c * (1.0 / p)
    d = b * l
    assert -b * l + d == 0.0
    l2 = d * (1.0 / b)
    assert -p * l2 + c != 0.0
    a = np.array([[p, b], [c, d]])
    assert O.inverse_branch(a, O.INVERSE_LU)[0] == O.SINGULAR_MATRIX
    rc_f, inv_f = O.inverse_branch(a, O.INVERSE_FULL_PIVOT)
    rc, inv, which = O.inverse(a)
    assert rc == O.OK and which == 1 and rc_f == O.OK
    assert np.array_equal(inv, inv_f) and np.all(np.isfinite(inv))
    assert _residual_ratio(a, inv) < 1e-14

    # 3. rank one, both LU forms meet an exact zero pivot, Householder QR
    #    leaves a rounding residue on R's diagonal and answers (nalgebra's
    #    QR try_inverse only rejects an exact zero diagonal).
    a = np.array([[1.0, 2.0], [2.0, 4.0]])
    assert -4.0 * (1.0 * 0.5) + 2.0 == 0.0  # LU: pivot 2, l = 1*(1/2)
    assert -2.0 * (2.0 * 0.25) + 1.0 == 0.0  # full pivot: 4 first, l = 2*(1/4)
    assert O.inverse_branch(a, O.INVERSE_LU)[0] == O.SINGULAR_MATRIX
    assert O.inverse_branch(a, O.INVERSE_FULL_PIVOT)[0] == O.SINGULAR_MATRIX
    rc, inv, which = O.inverse(a)
    assert rc == O.OK and which == 2
    assert np.array_equal(inv, O.inverse_branch(a, O.INVERSE_QR)[1])

    # 4. an exact zero column survives every elimination and reflection
    #    exactly: all three fail and the cascade reports SingularMatrix
    for a in (np.array([[1.0, 0.0], [2.0, 0.0]]), np.zeros((3, 3)),
              np.array([[2.0, 0.0, 1.0], [1.0, 0.0, 3.0], [5.0, 0.0, 4.0]])):
        for br in (O.INVERSE_LU, O.INVERSE_FULL_PIVOT, O.INVERSE_QR):
            assert O.inverse_branch(a, br)[0] == O.SINGULAR_MATRIX
        rc, _, which = O.inverse(a)
        assert rc == O.SINGULAR_MATRIX and which == -1


# ---------------------------------------------------------------- golden solutions
def test_lorenz_vs_dop853(orc):
    g = np.load("tests/golden/lorenz_golden.npz")
    for i in range(8):
        for k, t in enumerate(g["times"][:3]):
            r = O.solve(O.RK45, "lorenz", g["y0"][:, i], 0.0, t, 1e-12, 1e-12, 0.1,
                        params=g["params"][:, i], semantics=O.INTENDED)
            assert r["status"] == O.OK
            err = np.abs(r["y"] - g["y"][i, k]).max()
            assert err < 1e-7 * max(1.0, t), (i, t, err)
    # convergence: a looser tolerance is measurably worse
    i, k = 0, 1
    errs = []
    for tol in (1e-6, 1e-9, 1e-12):
        r = O.solve(O.RK45, "lorenz", g["y0"][:, i], 0.0, g["times"][k], tol, 1e-12, 0.1,
                    params=g["params"][:, i], semantics=O.INTENDED)
        errs.append(np.abs(r["y"] - g["y"][i, k]).max())
    assert errs[0] > errs[1] > errs[2]


def test_lv_vs_dop853_with_stops(orc):
    g = np.load("tests/golden/lv_golden.npz")
    obs = g["times"]
    for i in range(8):
        r = O.solve(O.RK45, "lotka_volterra", g["y0"][:, i], 0.0, obs[-1], 1e-11, 1e-12,
                    0.1, params=g["params"][:, i], semantics=O.INTENDED, stops=obs[:-1])
        assert r["status"] == O.OK
        got = np.vstack([r["stop_y"], r["y"][None, :]])
        assert np.abs(got - g["y"][i]).max() < 1e-8


def test_lv_invariant(orc):
    w = W.lotka_volterra(16)
    for i in range(16):
        a, b, g_, d = w["params"][:, i]
        y0 = w["y0"][:, i]
        r = O.solve(O.RK45, "lotka_volterra", y0, 0.0, 15.0, 1e-10, 1e-12, 0.1,
                    params=w["params"][:, i], semantics=O.INTENDED)
        V = lambda x, y: d * x - g_ * np.log(x) + b * y - a * np.log(y)  # noqa: E731
        assert abs(V(*r["y"]) - V(*y0)) < 1e-7


def test_linear_exact_solution(orc):
    w = W.linear_shared(4, N=128)
    A, bvec, Q, D = w["A"], w["b"], w["Q"], w["D"]
    shared = np.concatenate([A.reshape(-1), bvec])
    xs = np.linalg.solve(A, bvec)  # A^{-1} b
    for i in range(2):
        y0 = w["y0"][:, i]
        r = O.solve(O.RK45, "linear", y0, 0.0, 5.0, 1e-9, 1e-12, 0.1, shared=shared,
                    semantics=O.INTENDED)
        exact = Q @ (np.exp(-D * 5.0) * (Q.T @ (y0 + xs))) - xs
        assert r["status"] == O.OK
        assert np.abs(r["y"] - exact).max() < 1e-7


def test_brusselator_bdf6_vs_radau(orc):
    g = np.load("tests/golden/bruss_golden.npz")
    for i in range(2):
        r = O.solve(O.BDF6, "brusselator", g["y0"][:, i], 0.0, 1.0, 1e-7, 1e-12, 0.1,
                    params=g["params"][:, i], semantics=O.INTENDED)
        assert r["status"] == O.OK
        assert np.abs(r["y"] - g["y"][i, 0]).max() < 1e-4


@pytest.mark.parametrize("method", [O.BDF6, O.BDF2])
def test_robertson_bdf_vs_radau(orc, method):
    """BDF at dimension 3 on Robertson's stiff kinetics against scipy Radau
    (rtol 1e-12) at t = 10: within 1e-5 (tol 1e-6); pins the Robertson
    right-hand side and the dimension-generic BDF loop."""
    g = np.load("tests/golden/robertson_golden.npz")
    w = W.robertson(16)
    assert np.array_equal(w["params"], g["params"])
    o = O.solve_ensemble(method, "robertson", w["y0"], w["params"], w["t_end"], w["tol"],
                         dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.INTENDED)
    assert np.all(o["status"] == O.OK)
    assert np.abs(o["y"] - g["y"]).max() <= 1e-5


# ---------------------------------------------------------------- profiles
def test_device_profile_agrees_with_bacon_profile(orc):
    """The kernels' arithmetic (fused multiply-adds, the squared-norm
    controller with the polynomial root) changes results only at the
    tolerance's scale: on the plain Lorenz workload the two profiles' final
    states agree within 10x the requested tolerance."""
    w = W.lorenz_plain(128)
    outs = []
    for prof in (O.PROFILE_BACON, O.PROFILE_DEVICE):
        outs.append(O.solve_ensemble(O.RK45, "lorenz", w["y0"], w["params"], w["t_end"],
                                     w["tol"], dt_min=w["dt_min"], dt_max=w["dt_max"],
                                     semantics=O.INTENDED, profile=prof))
    assert np.all(outs[0]["status"] == O.OK) and np.all(outs[1]["status"] == O.OK)
    assert np.abs(outs[0]["y"] - outs[1]["y"]).max() < 10 * 1e-6


def test_builder_validation_order(orc):
    b = O.OracleBuilder()
    assert b.call("new", O.RK45, 0) == O.STATIC_ON_DYNAMIC
    assert b.call("new", O.RK45, 1) == O.OK
    assert b.call("with_tolerance", 0.0) == O.TOLERANCE_OOB
    assert b.call("with_maximum_dt", -1.0) == O.TIME_DELTA_OOB
    assert b.call("with_minimum_dt", 0.5) == O.OK
    assert b.call("with_maximum_dt", 0.1) == O.OK
    assert b.b.dt_min == 0.1  # lowered (rk.rs:184-189)
    assert b.call("with_ending_time", 1.0) == O.OK
    assert b.call("with_initial_time", 1.0) == O.TIME_START_OOB
    assert b.call("solve") == O.MISSING_PARAMETERS


# ---------------------------------------------------------------- Adams
ADAMS_CASES = [
    # method, rhs, y0, t_end, dt_min, dt_max, tol, exact    (adams.rs:696-922)
    ("adams5_exp", O.ADAMS5, "exp", 1.0, 2.0, 1e-5, 0.1, 0.0005, np.exp),
    ("adams5_quadratic", O.ADAMS5, "quadratic", 1.0, 5.0, 1e-7, 0.001, 0.01,
     lambda t: 1.0 - t * t),
    ("adams5_sine", O.ADAMS5, "cos", 0.0, 2 * np.pi, 1e-5, 0.001, 0.01, np.sin),
    ("adams3_exp", O.ADAMS3, "exp", 1.0, 2.0, 1e-5, 0.1, 0.001, np.exp),
    ("adams3_quadratic", O.ADAMS3, "quadratic", 1.0, 5.0, 1e-5, 0.001, 0.1,
     lambda t: 1.0 - t * t),
    ("adams3_sine", O.ADAMS3, "cos", 0.0, 2 * np.pi, 1e-5, 0.001, 0.01, np.sin),
]


@pytest.mark.parametrize("profile", [O.PROFILE_BACON, O.PROFILE_DEVICE])
@pytest.mark.parametrize("semantics", [O.AS_BUILT, O.INTENDED])
@pytest.mark.parametrize("case", ADAMS_CASES, ids=[c[0] for c in ADAMS_CASES])
def test_reference_adams(orc, case, semantics, profile):
    """The reference's Adams tests: every yielded point within 0.01 of the
    closed form (float_cmp approx_eq), and the solve reaches t_end."""
    name, method, rhs, y0, t_end, dt_min, dt_max, tol, exact = case
    r = O.solve(method, rhs, [y0], 0.0, t_end, tol, dt_min, dt_max,
                semantics=semantics, profile=profile, path_cap=200000)
    assert r["status"] == O.OK
    assert r["path_len"] > 0
    assert abs(r["t"] - t_end) < 1e-9
    for t, y in zip(r["path_t"], r["path_y"]):
        assert approx_eq(y[0], exact(t), 0.01)


@pytest.mark.parametrize("method,order", [(O.ADAMS5, 5), (O.ADAMS3, 3)])
def test_adams_coefficients_order_conditions(orc, method, order):
    """Predictor: Adams-Bashforth with (O-1) steps, weight pred[j] on
    f_{n-j}: sum_j pred[j] (-j)^q = 1/(q+1) for q < O-1.  Corrector:
    Adams-Moulton, corr[0] on f_{n+1}, corr[k] on f_{n+1-k}:
    sum_k corr[k] (1-k)^q = 1/(q+1) for q < O.  Error coefficient 19/270."""
    pred, corr, err = O.adams_coefficients(method)
    assert pred.size == order and corr.size == order
    for q in range(order - 1):
        s = sum(pred[j] * (-j) ** q for j in range(order - 1))
        assert abs(s - 1.0 / (q + 1)) < 1e-14, (q, s)
    assert pred[order - 1] == 0.0
    for q in range(order):
        s = sum(corr[k] * (1 - k) ** q for k in range(order))
        assert abs(s - 1.0 / (q + 1)) < 1e-14, (q, s)
    assert err == 19.0 / 270.0


@pytest.mark.parametrize("n", [3, 5])
def test_rootn_accuracy(orc, n):
    rng = np.random.default_rng(n)
    xs = np.concatenate([np.exp(rng.uniform(-690, 690, 4000)), [1.0, 2.0, 0.5, 7.0, 1e-300]])
    for x in xs:
        got = O.rootn(x, n)
        want = x ** (1.0 / n)
        assert abs(got / want - 1.0) < 3e-7, (x, got, want)
    assert O.rootn(0.0, n) == 0.0 and O.rootn(np.inf, n) == np.inf
    assert np.isnan(O.rootn(np.nan, n))


@pytest.mark.parametrize("profile", [O.PROFILE_BACON, O.PROFILE_DEVICE])
def test_stop_next_to_accepted_step(orc, profile):
    """A stop dt_min/8 after an accepted step time: the step cut short to land
    on it is accepted, and the next attempt gets the step size proposed
    before the cut, so the solve still succeeds (shrinking dt to the cut
    would fail it with MinimumTimeDeltaExceeded).  Up to that step the path
    is the stop-free path, bit for bit."""
    w = W.lorenz_plain(2)
    kw = dict(params=w["params"][:, 0], semantics=O.INTENDED, profile=profile, path_cap=200000)
    free = O.solve(O.RK45, "lorenz", w["y0"][:, 0], 0.0, 10.0, 1e-6, 1e-6, 0.1, **kw)
    assert free["status"] == O.OK
    k = 100
    stop = free["path_t"][k] + 1e-6 / 8
    r = O.solve(O.RK45, "lorenz", w["y0"][:, 0], 0.0, 10.0, 1e-6, 1e-6, 0.1,
                stops=np.array([stop]), **kw)
    assert r["status"] == O.OK and r["t"] == 10.0
    assert np.array_equal(r["path_t"][:k + 1], free["path_t"][:k + 1])
    assert np.array_equal(r["path_y"][:k + 1], free["path_y"][:k + 1])
    assert r["path_t"][k + 1] == stop
    # the step after the landing is the one the controller proposed at t_k
    # (a full step; the cut one was 1.25e-7 long)
    assert r["path_dt"][k + 2] == free["path_dt"][k + 1]


def test_root8_accuracy(orc):
    """The DEVICE controller's r^(1/8) against pow: relative error below
    1.11e-6 over the whole controller range (1e-8, 65536), a dense sweep of
    every binade and every residue of the exponent mod 8 (the 2^(k/8)
    scale table); exact at powers of 2^8; NaN in, NaN out."""
    rng = np.random.default_rng(8)
    xs = np.concatenate([np.exp(rng.uniform(np.log(1e-8), np.log(65536.0), 200000)),
                         2.0 ** np.arange(-26, 16, dtype=float),
                         np.nextafter(1.0, 0.0) ** np.arange(1, 4), [1.0, 1.5, 1.9999999]])
    worst = 0.0
    for x in xs:
        got = O.root8(x)
        worst = max(worst, abs(got / x ** 0.125 - 1.0))
    assert worst < 1.11e-6, worst
    assert worst > 1e-6  # the bound is the polynomial's, not a looser one
    for k in range(-3, 3):
        assert O.root8(2.0 ** (8 * k)) == 2.0 ** k * O.root8(1.0)
    assert np.isnan(O.root8(np.nan))


def test_adams_as_built_derivative_history_is_one_step_stale(orc):
    """As built, the first Adams step after an RK4 bootstrap is held back
    while the bootstrap values are yielded, and its derivative never joins
    the derivative deque (adams.rs:434-441 pushes the value only), so the
    next predictor and corrector use a history one step behind.  On the
    reference's scalar tests that still passes (above); on Lorenz-63 at tol
    1e-6 the error estimate then fails every step, dt collapses below dt_min
    and the solve ends with MinimumTimeDeltaExceeded near t = 0.32 after
    millions of RK4 bootstrap yields.  The intended variant (the derivative
    joins the history) finishes with 98 rejections."""
    kw = dict(t0=0.0, t_end=10.0, tol=1e-6, dt_min=1e-8, dt_max=0.1)
    a = O.solve(O.ADAMS5, "lorenz", [1.0, 1.0, 20.0], params=[10.0, 28.0, 8.0 / 3.0],
                semantics=O.AS_BUILT, **kw)
    b = O.solve(O.ADAMS5, "lorenz", [1.0, 1.0, 20.0], params=[10.0, 28.0, 8.0 / 3.0],
                semantics=O.INTENDED, **kw)
    assert a["status"] == O.MINIMUM_TIME_DELTA_EXCEEDED
    assert a["t"] < 0.5 and a["n_accepted"] > 10 ** 6
    assert b["status"] == O.OK and b["t"] == 10.0
    assert b["n_rejected"] < 200


def test_adams_lorenz_against_golden(orc):
    """Adams5 and Adams3 (intended) on golden Lorenz instances against the
    DOP853 1e-13 solution at t = 2: errors shrink with the tolerance, and
    Adams5 at tol 1e-10 is within 1e-9."""
    g = np.load("tests/golden/lorenz_golden.npz")
    # Adams3 (2nd order) is held to tol 1e-8: at 1e-10 it needs ~10^7 steps
    for method, tols in ((O.ADAMS5, (1e-6, 1e-8, 1e-10)), (O.ADAMS3, (1e-6, 1e-8))):
        for i in range(4):
            errs = []
            for tol in tols:
                r = O.solve(method, "lorenz", g["y0"][:, i], 0.0, g["times"][2], tol, 1e-12,
                            0.1, params=g["params"][:, i], semantics=O.INTENDED)
                assert r["status"] == O.OK
                errs.append(np.abs(r["y"] - g["y"][i, 2]).max())
            assert all(a > b for a, b in zip(errs, errs[1:])), (method, i, errs)
            if method == O.ADAMS5:
                assert errs[-1] < 1e-9, (i, errs)


# ---- complex state (N: ComplexField, rk.rs:53, adams.rs:44) ------------------

def test_complex_norm_groups_each_component():
    """nalgebra's dotc over complex entries adds re^2 + im^2 as one term per
    entry.  With three entries (1, 0), (a, a), (a, a), a^2 = 0.72 half-ulps
    of 1: summing the six reals one at a time never leaves 1, while each
    complex term (1.44 half-ulps) rounds up, so only the grouped norm exceeds
    1.  The device profile is the sequential fused sum of the reals."""
    a = 1.2 * 2.0 ** -27
    x = np.array([1.0, 0.0, a, a, a, a])
    assert O.norm(x) == 1.0
    assert O.norm(x, profile=O.PROFILE_DEVICE) == 1.0
    assert O.norm(x, complex_pairs=True) == np.nextafter(1.0, 2.0)
    # away from rounding ties all three agree with numpy's complex norm
    rng = np.random.default_rng(3)
    for _ in range(20):
        z = rng.standard_normal(4) + 1j * rng.standard_normal(4)
        xr = np.stack([z.real, z.imag], 1).ravel()
        ref = np.linalg.norm(z)
        for kw in ({}, {"complex_pairs": True}, {"profile": O.PROFILE_DEVICE}):
            assert abs(O.norm(xr, **kw) - ref) <= 4e-16 * ref


def test_two_level_rhs_is_the_complex_product():
    """The registered two-level RHS equals (-i H - gamma) c computed in
    complex arithmetic, in both profiles, to rounding."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        hd, ho = rng.uniform(-2, 2, 2)
        g = rng.uniform(0.01, 0.2)
        y = rng.standard_normal(4)
        c = np.array([y[0] + 1j * y[1], y[2] + 1j * y[3]])
        H = np.array([[hd, ho], [ho, -hd]])
        f = -1j * (H @ c) - g * c
        want = np.array([f[0].real, f[0].imag, f[1].real, f[1].imag])
        for prof in (O.PROFILE_BACON, O.PROFILE_DEVICE):
            rc, got = O.rhs_eval("two_level", y, params=[hd, ho, g], profile=prof)
            assert rc == O.OK
            assert np.allclose(got, want, rtol=1e-14, atol=1e-14 * np.abs(want).max())


def test_two_level_against_closed_form():
    """RK45, RK23 and Adams5 (both profiles) on 32 complex two-level systems
    against c(t) = e^{-gamma t} exp(-iHt) c0 at t = 10, tol 1e-8: within
    2e-6 (20 tol t_end), and the squared modulus decays as e^{-2 gamma t}."""
    w = W.two_level(32)
    ex = W.two_level_exact(w["params"], w["y0"], 10.0)
    for method in (O.RK45, O.RK23, O.ADAMS5):
        for prof in (O.PROFILE_BACON, O.PROFILE_DEVICE):
            for i in range(32):
                r = O.solve(method, "two_level", w["y0"][:, i], 0.0, 10.0, 1e-8, 1e-12, 0.1,
                            params=w["params"][:, i], semantics=O.INTENDED, profile=prof)
                assert r["status"] == O.OK and r["t"] == 10.0
                assert np.abs(r["y"] - ex[:, i]).max() < 2e-6, (method, prof, i)
                decay = np.exp(-2 * w["params"][2, i] * 10.0)
                assert abs((r["y"] ** 2).sum() - decay) < 2e-6


def test_bdf_rejects_complex_state():
    """BDF's Broyden update uses the plain transpose (bdf.rs:461-466), so a
    complex system is not its real-pair system there: the oracle refuses."""
    w = W.two_level(1)
    with pytest.raises(ValueError):
        O.solve(O.BDF6, "two_level", w["y0"][:, 0], 0.0, 1.0, 1e-8, 1e-12, 0.1,
                params=w["params"][:, 0], semantics=O.INTENDED)
