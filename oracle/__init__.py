"""ctypes front end of the CPU oracle (bacon_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates bacon-sci 0.16.2's adaptive IVP
step loop (src/ivp/rk.rs:249-423, src/ivp/bdf.rs:257-634) in C.  It is imported
by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
cpu_baseline leg as the timed CPU path.  The product (bacon_amd/) never imports
it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# The checker build (portable -O2); BACON_ORACLE_LIB may name another build of
# the same source (the sanitizer build, liboracle_san.so).
CHECKER_PATH = os.path.join(HERE, "liboracle.so")
LIB_PATH = os.environ.get("BACON_ORACLE_LIB") or CHECKER_PATH

# status codes (IVPError declaration order, src/ivp.rs:50-76)
OK = 0
MISSING_PARAMETERS = 1
USER_ERROR = 2
TOLERANCE_OOB = 3
TIME_DELTA_OOB = 4
TIME_END_OOB = 5
TIME_START_OOB = 6
FROM_PRIMITIVE_FAILURE = 7
MINIMUM_TIME_DELTA_EXCEEDED = 8
MAXIMUM_ITERATIONS_EXCEEDED = 9
SINGULAR_MATRIX = 10
DYNAMIC_ON_STATIC = 11
STATIC_ON_DYNAMIC = 12
INVALID_ARGUMENT = -1

RK45, RK23, BDF6, BDF2, ADAMS5, ADAMS3 = 0, 1, 2, 3, 4, 5
AS_BUILT, INTENDED = 0, 1
PROFILE_BACON, PROFILE_DEVICE = 0, 1

RHS = {
    "exp": 0, "neg": 1, "quadratic": 2, "cos": 3, "lorenz": 4, "linear": 5,
    "brusselator": 6, "lotka_volterra": 7, "oscillator": 8, "fail_after": 9,
    "lv_sens": 10, "robertson": 11, "two_level": 12,
}

_dp = C.POINTER(C.c_double)


class Problem(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("semantics", C.c_int32), ("profile", C.c_int32),
        ("rhs", C.c_int32), ("dim", C.c_int32), ("n_stops", C.c_int32),
        ("t0", C.c_double), ("t_end", C.c_double), ("tol", C.c_double),
        ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("max_attempts", C.c_int64),
        ("y0", _dp), ("params", _dp), ("shared", _dp), ("stops", _dp),
    ]


class Result(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("status", C.c_int32),
        ("pad", C.c_int32), ("n_accepted", C.c_int64), ("n_rejected", C.c_int64),
        ("n_rhs", C.c_int64), ("n_newton", C.c_int64), ("n_lu", C.c_int64),
        ("n_yield", C.c_int64),
    ]


class Path(C.Structure):
    _fields_ = [("cap", C.c_int64), ("len", C.c_int64),
                ("t", _dp), ("y", _dp), ("dt", _dp)]


class Builder(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "method", "is_dyn", "dim", "built", "have_tol", "have_dt_max",
        "have_dt_min", "have_t0", "have_t_end", "have_y0", "have_deriv")] + [
        (n, C.c_double) for n in ("tol", "dt_max", "dt_min", "t0", "t_end")]


def build(force: bool = False, target: str = "liboracle.so") -> str:
    """Compile a Makefile target (the checker liboracle.so by default; the
    CPU baseline's liboracle_native.so; the sanitizer build liboracle_san.so)."""
    path = os.path.join(HERE, target)
    if force or not os.path.exists(path):
        cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else []) + [target]
        subprocess.run(cmd, check=True)
    return path


def use_native_build() -> str:
    """Switch this process to the -O3 -march=native build, compiled here for
    this host (the CPU baseline's timing); call before the first solve."""
    global LIB_PATH, _lib
    LIB_PATH = build(force=True, target="liboracle_native.so")
    _lib = None
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH) and LIB_PATH == CHECKER_PATH:
            build()
        L = C.CDLL(LIB_PATH)
        L.or_solve.argtypes = [C.POINTER(Problem), _dp, C.POINTER(Result),
                               C.POINTER(Path), _dp]
        L.or_solve.restype = C.c_int32
        L.or_solve_ensemble.argtypes = [
            C.POINTER(Problem), C.c_int64, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp,
            C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int32]
        L.or_solve_ensemble.restype = C.c_int32
        L.or_rhs_eval.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_double, _dp,
                                  _dp, _dp, _dp]
        L.or_rhs_eval.restype = C.c_int32
        L.or_inverse.argtypes = [C.c_int32, _dp, _dp, C.POINTER(C.c_int32)]
        L.or_lu_inverse.argtypes = [C.c_int32, _dp, _dp]
        L.or_inverse_branch.argtypes = [C.c_int32, _dp, _dp, C.c_int32]
        L.or_rk_tableau.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp, _dp]
        L.or_bdf_coefficients.argtypes = [C.c_int32, _dp, _dp]
        L.or_adams_coefficients.argtypes = [C.c_int32, _dp, _dp, _dp]
        L.or_rootn.argtypes = [C.c_double, C.c_int32]
        L.or_rootn.restype = C.c_double
        L.or_root8.argtypes = [C.c_double]
        L.or_root8.restype = C.c_double
        L.or_norm.argtypes = [C.c_int32, _dp, C.c_int32, C.c_int32]
        L.or_norm.restype = C.c_double
        L.or_rhs_is_complex.argtypes = [C.c_int32]
        L.or_rk_replay.argtypes = [C.POINTER(Problem), C.c_int64, _dp, _dp, _dp]
        for name in ("or_builder_new", "or_builder_new_dyn"):
            getattr(L, name).argtypes = [C.POINTER(Builder), C.c_int32, C.c_int32]
        for name in ("or_builder_with_tolerance", "or_builder_with_maximum_dt",
                     "or_builder_with_minimum_dt", "or_builder_with_initial_time",
                     "or_builder_with_ending_time"):
            getattr(L, name).argtypes = [C.POINTER(Builder), C.c_double]
        L.or_builder_with_initial_conditions.argtypes = [C.POINTER(Builder), C.c_int32]
        L.or_builder_with_derivative.argtypes = [C.POINTER(Builder)]
        L.or_builder_solve.argtypes = [C.POINTER(Builder)]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def solve(method=RK45, rhs="lorenz", y0=(1.0,), t0=0.0, t_end=1.0, tol=1e-6,
          dt_min=1e-6, dt_max=0.1, params=None, shared=None,
          semantics=INTENDED, profile=PROFILE_BACON, path_cap=0, stops=None,
          max_attempts=0):
    """Solve one IVP.  Returns a dict with t, y, status, counters and, when
    path_cap > 0, the yielded path (t, y, dt)."""
    y0 = _f64(y0)
    params = _f64(params if params is not None else [0.0])
    shared = _f64(shared)
    stops = _f64(stops)
    dim = y0.size
    p = Problem(method=method, semantics=semantics, profile=profile,
                rhs=RHS[rhs] if isinstance(rhs, str) else rhs, dim=dim,
                n_stops=0 if stops is None else stops.size,
                t0=t0, t_end=t_end, tol=tol, dt_min=dt_min, dt_max=dt_max,
                max_attempts=max_attempts, y0=_ptr(y0), params=_ptr(params),
                shared=_ptr(shared), stops=_ptr(stops))
    y_out = np.zeros(dim)
    r = Result()
    path = None
    if path_cap:
        pt = np.zeros(path_cap)
        py = np.zeros((path_cap, dim))
        pdt = np.zeros(path_cap)
        path = Path(cap=path_cap, len=0, t=_ptr(pt), y=_ptr(py), dt=_ptr(pdt))
    stop_y = None if stops is None else np.zeros((stops.size, dim))
    rc = lib().or_solve(C.byref(p), _ptr(y_out), C.byref(r),
                        C.byref(path) if path is not None else None, _ptr(stop_y))
    if rc == -1:
        raise ValueError("or_solve: invalid argument (method, dimension, stops or "
                         "right-hand side not supported together)")
    out = dict(t=r.t, dt=r.dt, y=y_out, status=r.status, n_accepted=r.n_accepted,
               n_rejected=r.n_rejected, n_rhs=r.n_rhs, n_newton=r.n_newton,
               n_lu=r.n_lu, n_yield=r.n_yield, stop_y=stop_y)
    if path is not None:
        n = min(path.len, path_cap)
        out.update(path_len=path.len, path_t=pt[:n], path_y=py[:n], path_dt=pdt[:n])
    return out


def solve_ensemble(method, rhs, y0_soa, params_soa, t_end, tol, t0=0.0,
                   dt_min=1e-6, dt_max=0.1, t_end_scalar=None, tol_scalar=None,
                   shared=None, semantics=INTENDED, profile=PROFILE_BACON,
                   threads=0, max_attempts=0):
    """Ensemble solve with structure-of-arrays inputs: y0_soa[dim, n],
    params_soa[n_params, n]; t_end / tol per instance arrays or None."""
    y0_soa = _f64(y0_soa)
    dim, n = y0_soa.shape
    params_soa = _f64(params_soa) if params_soa is not None else np.zeros((0, n))
    shared = _f64(shared)
    t_end = _f64(t_end)
    tol = _f64(tol)
    dummy_y0 = np.zeros(dim)
    p = Problem(method=method, semantics=semantics, profile=profile,
                rhs=RHS[rhs] if isinstance(rhs, str) else rhs, dim=dim, n_stops=0,
                t0=t0, t_end=t_end_scalar if t_end_scalar is not None else 1.0,
                tol=tol_scalar if tol_scalar is not None else 1e-6,
                dt_min=dt_min, dt_max=dt_max, max_attempts=max_attempts,
                y0=_ptr(dummy_y0), params=None, shared=_ptr(shared), stops=None)
    t_out = np.zeros(n)
    y_out = np.zeros((dim, n))
    status = np.zeros(n, dtype=np.int32)
    nacc = np.zeros(n, dtype=np.int64)
    nrej = np.zeros(n, dtype=np.int64)
    lib().or_solve_ensemble(
        C.byref(p), n, params_soa.shape[0], _ptr(y0_soa), _ptr(params_soa),
        _ptr(t_end), _ptr(tol), _ptr(t_out), _ptr(y_out),
        status.ctypes.data_as(C.POINTER(C.c_int32)),
        nacc.ctypes.data_as(C.POINTER(C.c_int64)),
        nrej.ctypes.data_as(C.POINTER(C.c_int64)), threads)
    return dict(t=t_out, y=y_out, status=status, n_accepted=nacc, n_rejected=nrej)


def rhs_eval(rhs, y, t=0.0, params=None, shared=None, profile=PROFILE_BACON):
    y = _f64(y)
    params = _f64(params if params is not None else [0.0])
    shared = _f64(shared)
    out = np.zeros_like(y)
    rc = lib().or_rhs_eval(RHS[rhs] if isinstance(rhs, str) else rhs, profile, y.size,
                           t, _ptr(y), _ptr(params), _ptr(shared), _ptr(out))
    return rc, out


def inverse(a):
    """nalgebra-semantics inverse cascade (LU, full-pivot LU, QR)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    acm = np.asfortranarray(a)
    inv = np.zeros((n, n), order="F")
    which = C.c_int32(-1)
    rc = lib().or_inverse(n, acm.ctypes.data_as(_dp), inv.ctypes.data_as(_dp),
                          C.byref(which))
    return rc, np.asarray(inv), which.value


INVERSE_LU, INVERSE_FULL_PIVOT, INVERSE_QR = 0, 1, 2


def inverse_branch(a, branch):
    """One member of the cascade alone: INVERSE_LU, INVERSE_FULL_PIVOT or
    INVERSE_QR.  Returns (rc, inverse)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    acm = np.asfortranarray(a)
    inv = np.zeros((n, n), order="F")
    rc = lib().or_inverse_branch(n, acm.ctypes.data_as(_dp), inv.ctypes.data_as(_dp), branch)
    return rc, np.asarray(inv)


def rk_tableau(method=RK45, semantics=INTENDED):
    c = np.zeros(6); m = np.zeros(36); b = np.zeros(6); e = np.zeros(6)
    s = lib().or_rk_tableau(method, semantics, _ptr(c), _ptr(m), _ptr(b), _ptr(e))
    return c[:s], m[:s * s].reshape(s, s), b[:s], e[:s]


def bdf_coefficients(method=BDF6):
    hi = np.zeros(7); lo = np.zeros(7)
    o = lib().or_bdf_coefficients(method, _ptr(hi), _ptr(lo))
    return hi[:o], lo[:o]


def adams_coefficients(method=ADAMS5):
    pred = np.zeros(5); corr = np.zeros(5); err = np.zeros(1)
    o = lib().or_adams_coefficients(method, _ptr(pred), _ptr(corr), _ptr(err))
    return pred[:o], corr[:o], float(err[0])


def rootn(x, n):
    return lib().or_rootn(float(x), n)


def root8(r):
    """The DEVICE profile's step-factor root r^(1/8), before the tangent bound."""
    return lib().or_root8(float(r))


def norm(x, profile=PROFILE_BACON, complex_pairs=False):
    """The error norm of a step: nalgebra's norm() (BACON; over complex
    (re, im) pairs with complex_pairs) or the device's fused sum (DEVICE)."""
    x = _f64(x)
    return lib().or_norm(x.size, _ptr(x), profile, int(complex_pairs))


def rk_replay(rhs, y0, t_seq, dt_seq, params=None, method=RK45,
              semantics=INTENDED, profile=PROFILE_BACON):
    """Frozen-step replay of the RK step map; returns the state after each step."""
    y0 = _f64(y0)
    params = _f64(params if params is not None else [0.0])
    t_seq = _f64(t_seq)
    dt_seq = _f64(dt_seq)
    n = t_seq.size
    p = Problem(method=method, semantics=semantics, profile=profile,
                rhs=RHS[rhs] if isinstance(rhs, str) else rhs, dim=y0.size,
                y0=_ptr(y0), params=_ptr(params))
    out = np.zeros((n, y0.size))
    rc = lib().or_rk_replay(C.byref(p), n, _ptr(t_seq), _ptr(dt_seq), _ptr(out))
    return rc, out


class OracleBuilder:
    """Builder restatement used to cross-check the C-ABI builder's validation."""

    def __init__(self):
        self.b = Builder()

    def call(self, name, *args):
        return getattr(lib(), "or_builder_" + name)(C.byref(self.b), *args)
