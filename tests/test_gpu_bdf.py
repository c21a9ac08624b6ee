"""GPU parity of the batched BDF kernels (through the C ABI) against the oracle.

Bars:
  * the reference's own BDF tests (src/ivp/bdf.rs:785-1063) through the
    single-instance builder and iterator: as built, the path is empty (the
    first BDF step is rejected and the time restore of bdf.rs:622 ends the
    solve); intended, every yielded value passes the test's own check, and
    the path has the oracle's length with every yielded (t, y) within 10x the
    tolerance of the oracle's;
  * the 32-component Brusselator ensemble: final states within 10x the
    requested tolerance (1e-7) of the oracle for each of 2048 instances, with
    the Jacobian inverse in FP64 and in FP32; as built, every instance fails
    with MaximumIterationsExceeded at t = 0.35 like the oracle;
  * SingularMatrix (bdf.rs:441): inputs whose finite-difference Jacobian is
    exactly zero, so LU, full-pivot LU and QR all fail, give status 10 at the
    oracle's time on the single-instance builder and the ensemble.
The BDF kernels reduce across lanes in a different order from the oracle, so
the comparison is to the tolerance, not to the bit.
"""
import numpy as np
import pytest

import oracle as O
from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

BDF_CASES = [
    ("exp", "exp", 1.0, (7.0, 7.0), np.exp),
    ("unstable", "neg", 1.0, (10.0, 10.0), lambda t: np.exp(-t)),
    ("quadratic", "quadratic", 1.0, (2.0, 1.0), lambda t: 1.0 - t * t),
    ("sin", "cos", 0.0, (6.0, 6.0), np.sin),
]


def _solver(cls, rhs, y0, t_end, semantics):
    return (cls.new(1).with_minimum_dt(1e-5).with_maximum_dt(0.1).with_tolerance(0.00001)
            .with_initial_time(0.0).with_ending_time(t_end)
            .with_initial_conditions_slice([y0]).with_derivative(rhs)
            .with_semantics(semantics))


@pytest.mark.parametrize("case", BDF_CASES, ids=[c[0] for c in BDF_CASES])
@pytest.mark.parametrize("cls", [ivp.BDF6, ivp.BDF2])
def test_reference_bdf_as_built_empty_path(gpu, orc, cls, case):
    name, rhs, y0, tends, exact = case
    t_end = tends[0] if cls is ivp.BDF6 else tends[1]
    path = _solver(cls, rhs, y0, t_end, L.AS_BUILT).solve().collect_vec()
    assert path == []


@pytest.mark.parametrize("case", BDF_CASES, ids=[c[0] for c in BDF_CASES])
@pytest.mark.parametrize("cls", [ivp.BDF6, ivp.BDF2])
def test_reference_bdf_intended(gpu, orc, cls, case):
    name, rhs, y0, tends, exact = case
    t_end = tends[0] if cls is ivp.BDF6 else tends[1]
    path = _solver(cls, rhs, y0, t_end, L.INTENDED).solve().collect_vec()
    assert len(path) > 0
    for t, y in path:
        assert abs(y[0] - exact(t)) <= 0.01
    method = O.BDF6 if cls is ivp.BDF6 else O.BDF2
    r = O.solve(method, rhs, [y0], 0.0, t_end, 1e-5, 1e-5, 0.1, semantics=O.INTENDED,
                path_cap=100000)
    assert r["path_len"] > 0
    assert len(path) == r["path_len"]
    tg = np.array([p[0] for p in path])
    yg = np.array([p[1][0] for p in path])
    assert np.all(np.abs(tg - r["path_t"]) <= 1e-9 * np.maximum(1.0, np.abs(r["path_t"])))
    yo = r["path_y"][:, 0]
    assert np.all(np.abs(yg - yo) <= 10 * 1e-5 * np.maximum(1.0, np.abs(yo)))


def test_bdf_user_error(gpu, orc):
    s = (ivp.BDF6.new(1).with_minimum_dt(1e-6).with_maximum_dt(0.05).with_tolerance(1e-6)
         .with_initial_time(0.0).with_ending_time(2.0).with_initial_conditions([1.0])
         .with_derivative("fail_after", [0.5]))
    with pytest.raises(ivp.IVPError) as e:
        s.solve().collect_vec()
    assert e.value.name == "UserError"


def _bruss(n):
    w = W.brusselator(n)
    w["dt_min"] = 1e-12
    w["dt_max"] = 0.1
    return w


@pytest.fixture(scope="module")
def bruss_oracle():
    w = _bruss(2048)
    o = O.solve_ensemble(O.BDF6, "brusselator", w["y0"], w["params"], w["t_end"], w["tol"],
                         dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.INTENDED)
    return w, o


@pytest.mark.parametrize("lu_bits", [64, 32])
def test_brusselator_ensemble_vs_oracle(gpu, orc, bruss_oracle, lu_bits):
    w, o = bruss_oracle
    g = ivp.solve_ensemble(L.BDF6, w, lu_bits=lu_bits)
    assert np.all(g["status"] == 0)
    assert np.all(np.abs(g["t"] - 10.0) < 1e-9)
    err = np.abs(g["y"] - o["y"]).max()
    assert err < 10 * 1e-7, err
    # accepted-step counts follow the oracle closely
    assert np.abs(g["n_accepted"] - o["n_accepted"]).max() <= 0.05 * o["n_accepted"].max()
    assert np.all(g["n_lu"] > 0) and np.all(g["n_newton"] > 0)


def test_brusselator_bdf2_vs_oracle(gpu, orc):
    """BDF2 (bdf.rs:708-762) on the same kernel: 8 Brusselator instances to
    t = 2, within 10 tol of the oracle, accepted counts within 5 %."""
    w = _bruss(8)
    w["t_end"] = np.full(8, 2.0)
    o = O.solve_ensemble(O.BDF2, "brusselator", w["y0"], w["params"], w["t_end"], w["tol"],
                         dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.INTENDED)
    g = ivp.solve_ensemble(L.BDF2, w)
    assert np.all(o["status"] == 0) and np.all(g["status"] == 0)
    assert np.all(np.abs(g["t"] - 2.0) < 1e-9)
    err = np.abs(g["y"] - o["y"]).max()
    assert err < 10 * 1e-7, err
    assert np.abs(g["n_accepted"] - o["n_accepted"]).max() <= 0.05 * o["n_accepted"].max()


def test_brusselator_vs_radau(gpu, orc):
    gold = np.load("tests/golden/bruss_golden.npz")
    w = _bruss(8)
    g = ivp.solve_ensemble(L.BDF6, w)
    assert np.array_equal(w["y0"], gold["y0"])
    assert np.abs(g["y"].T - gold["y"][:, 1]).max() < 1e-4


def test_brusselator_as_built_fails_like_oracle(gpu, orc):
    w = _bruss(16)
    g = ivp.solve_ensemble(L.BDF6, w, semantics=L.AS_BUILT)
    o = O.solve_ensemble(O.BDF6, "brusselator", w["y0"], w["params"], w["t_end"], w["tol"],
                         dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.AS_BUILT)
    assert np.all(o["status"] == 9)
    assert np.array_equal(g["status"], o["status"])
    assert np.allclose(g["t"], o["t"])


# ------------------------------------------------------------- SingularMatrix
@pytest.mark.parametrize("cls,method", [(ivp.BDF6, O.BDF6), (ivp.BDF2, O.BDF2)])
def test_singular_matrix_scalar(gpu, orc, cls, method):
    """As built, y' = y from y0 = 0: the RK4 bootstrap keeps y = 0, and the
    finite-difference column (G(dt) + G(-dt)) / (2 dt) (bdf.rs:407) is
    exactly 0 (G is odd), so lu, full_piv_lu and qr all fail to invert it and
    the solve returns SingularMatrix (bdf.rs:441) at the end of the
    bootstrap."""
    r = O.solve(method, "exp", [0.0], 0.0, 1.0, 1e-4, 1e-6, 0.1, semantics=O.AS_BUILT,
                path_cap=100)
    assert r["status"] == O.SINGULAR_MATRIX and r["path_len"] == 0
    s = (cls.new(1).with_minimum_dt(1e-6).with_maximum_dt(0.1).with_tolerance(1e-4)
         .with_initial_time(0.0).with_ending_time(1.0).with_initial_conditions_slice([0.0])
         .with_derivative("exp").with_semantics(L.AS_BUILT))
    with pytest.raises(ivp.IVPError) as e:
        s.solve().collect_vec()
    assert e.value.name == "SingularMatrix"
    n = 96
    w = dict(rhs="exp", dim=1, y0=np.zeros((1, n)), params=np.zeros((0, n)), t0=0.0,
             t_end=np.full(n, 1.0), tol=np.full(n, 1e-4), dt_min=1e-6, dt_max=0.1)
    g = ivp.solve_ensemble(cls.METHOD, w, semantics=L.AS_BUILT)
    assert np.all(g["status"] == L.SINGULAR_MATRIX)
    assert np.all(g["t"] == r["t"])
    assert np.all(g["y"] == 0.0)


@pytest.mark.parametrize("lu_bits", [64, 32])
@pytest.mark.parametrize("method", [L.BDF6, L.BDF2])
def test_singular_matrix_brusselator(gpu, orc, method, lu_bits):
    """The 32-component path (one instance per wave, split rows): the
    Brusselator with A = B = alpha = 0 from y = 0 stays at 0 through the
    bootstrap and, as built, every finite-difference column is exactly 0, so
    the pivot search finds no non-zero pivot and the solve returns
    SingularMatrix, like the oracle's LU / full-pivot LU / QR cascade."""
    om = O.BDF6 if method == L.BDF6 else O.BDF2
    r = O.solve(om, "brusselator", np.zeros(32), 0.0, 1.0, 1e-7, 1e-12, 0.1,
                params=[0.0, 0.0, 0.0], semantics=O.AS_BUILT)
    assert r["status"] == O.SINGULAR_MATRIX
    n = 8
    w = _bruss(n)
    w["y0"] = np.zeros((32, n))
    w["params"] = np.zeros((3, n))
    w["t_end"] = np.full(n, 1.0)
    g = ivp.solve_ensemble(method, w, semantics=L.AS_BUILT, lu_bits=lu_bits)
    assert np.all(g["status"] == L.SINGULAR_MATRIX)
    assert np.all(g["t"] == r["t"])
    # intended: the central difference is the identity-like (1 + dt b0) I,
    # so the same input is not singular (the Broyden update then divides
    # 0 by 0 on the zero residual and fails like the oracle, bdf.rs:462-466)
    ri = O.solve(om, "brusselator", np.zeros(32), 0.0, 1.0, 1e-7, 1e-12, 0.1,
                 params=[0.0, 0.0, 0.0], semantics=O.INTENDED)
    gi = ivp.solve_ensemble(method, w, semantics=L.INTENDED, lu_bits=lu_bits)
    assert ri["status"] == O.MAXIMUM_ITERATIONS_EXCEEDED
    assert np.all(gi["status"] == ri["status"])
