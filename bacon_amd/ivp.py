"""Host-side mirror of bacon-sci's `ivp` builder API over the C ABI.

Names, argument meaning and error behaviour follow src/ivp.rs:134-239 and
src/ivp/rk.rs:136-256 (src/ivp/bdf.rs:144-264 is identical):

    path = (RungeKutta45.new(1)
            .with_minimum_dt(1e-4).with_maximum_dt(0.1)
            .with_initial_time(0.0).with_ending_time(10.0)
            .with_tolerance(1e-5)
            .with_initial_conditions_slice([1.0])
            .with_derivative("quadratic")
            .solve()
            .collect_vec())

Every setter validates eagerly and raises IVPError (the Result::Err of the
reference); solve() raises MissingParameters; iteration yields (t, y) and
raises the solver's failure once, after which the iterator is exhausted.
The README spellings (with_dt_min, with_dt_max, with_start, with_end) are
aliases.  The right-hand side is a registered device function named by id,
with per-instance parameters standing in for the reference's UserData.

`Ensemble` is the batched builder: the same setters taking per-instance
arrays (or scalars), solved by one persistent GPU launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

ERROR_NAMES = {
    1: "MissingParameters", 2: "UserError", 3: "ToleranceOOB", 4: "TimeDeltaOOB",
    5: "TimeEndOOB", 6: "TimeStartOOB", 7: "FromPrimitiveFailure",
    8: "MinimumTimeDeltaExceeded", 9: "MaximumIterationsExceeded",
    10: "SingularMatrix", 11: "DynamicOnStatic", 12: "StaticOnDynamic",
    -1: "InvalidArgument", -2: "HipError", -3: "Unsupported", -4: "StateError",
}


class IVPError(Exception):
    """IVPError (src/ivp.rs:50-76); `.code` is the C ABI status."""

    def __init__(self, code: int):
        self.code = int(code)
        self.name = ERROR_NAMES.get(self.code, f"status {code}")
        super().__init__(self.name)


def _check(rc: int):
    if rc != L.OK:
        raise IVPError(rc)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _rhs_id(rhs) -> int:
    return L.RHS[rhs] if isinstance(rhs, str) else int(rhs)


class IVPIterator:
    """IVPIterator (src/ivp.rs:192-239) over a path produced on the GPU."""

    def __init__(self, solver: "_Solver"):
        self._s = solver

    def __iter__(self):
        return self

    def __next__(self):
        dim = self._s.dim()
        t = C.c_double()
        y = np.zeros(dim)
        rc = L.lib().bg_next(self._s._h, C.byref(t), _dptr(y))
        if rc == L.OK:
            return t.value, y
        if rc == L.DONE:
            raise StopIteration
        raise IVPError(rc)

    def collect_vec(self):
        """Path::collect_vec (src/ivp.rs:209-211): all values, or the error."""
        return list(self)

    def time(self) -> float:
        t = C.c_double()
        _check(L.lib().bg_time(self._s._h, C.byref(t)))
        return t.value


class _Solver:
    METHOD = L.RK45

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().bg_destroy(h)
            except Exception:
                pass
            self._h = None

    @classmethod
    def new(cls, dim: int = 1, dim_kind: int = L.DIM_STATIC):
        """new() for a solver with a static dimension `dim` (rk.rs:136-148)."""
        h = C.c_void_p()
        _check(L.lib().bg_new(cls.METHOD, dim_kind, dim, C.byref(h)))
        return cls(h)

    @classmethod
    def new_dyn(cls, size: int, dim_kind: int = L.DIM_DYN):
        """new_dyn(size) (rk.rs:150-162)."""
        h = C.c_void_p()
        _check(L.lib().bg_new_dyn(cls.METHOD, dim_kind, size, C.byref(h)))
        return cls(h)

    def dim(self) -> int:
        d = C.c_uint32()
        _check(L.lib().bg_dim(self._h, C.byref(d)))
        return d.value

    def _set(self, fn, *args):
        _check(getattr(L.lib(), fn)(self._h, *args))
        return self

    def with_tolerance(self, tol):
        return self._set("bg_with_tolerance", float(tol))

    def with_maximum_dt(self, v):
        return self._set("bg_with_maximum_dt", float(v))

    def with_minimum_dt(self, v):
        return self._set("bg_with_minimum_dt", float(v))

    def with_initial_time(self, v):
        return self._set("bg_with_initial_time", float(v))

    def with_ending_time(self, v):
        return self._set("bg_with_ending_time", float(v))

    def with_initial_conditions_slice(self, y0):
        y = np.ascontiguousarray(y0, dtype=np.float64)
        self._y0_keep = y
        return self._set("bg_with_initial_conditions_slice", _dptr(y), y.size)

    with_initial_conditions = with_initial_conditions_slice

    def with_derivative(self, rhs, params=()):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._p_keep = p
        return self._set("bg_with_derivative", _rhs_id(rhs), _dptr(p), p.size)

    def with_semantics(self, semantics: int):
        return self._set("bg_with_semantics", semantics)

    # README.md:33-38 spellings
    def with_dt_min(self, v):
        return self._set("bg_with_dt_min", float(v))

    def with_dt_max(self, v):
        return self._set("bg_with_dt_max", float(v))

    def with_start(self, v):
        return self._set("bg_with_start", float(v))

    def with_end(self, v):
        return self._set("bg_with_end", float(v))

    def solve(self, data=None) -> IVPIterator:
        """solve(data) (rk.rs:249-343); `data` is unused (parameters travel
        with with_derivative)."""
        _check(L.lib().bg_solve(self._h))
        return IVPIterator(self)


class RungeKutta45(_Solver):
    """RungeKutta45 (rk.rs:561)."""
    METHOD = L.RK45


class RungeKutta23(_Solver):
    """RungeKutta23 (rk.rs:656)."""
    METHOD = L.RK23


class BDF6(_Solver):
    """BDF6 (bdf.rs:706)."""
    METHOD = L.BDF6


class BDF2(_Solver):
    """BDF2 (bdf.rs:762)."""
    METHOD = L.BDF2


class Adams5(_Solver):
    """Adams5 (adams.rs:633)."""
    METHOD = L.ADAMS5


class Adams3(_Solver):
    """Adams3 (adams.rs:693)."""
    METHOD = L.ADAMS3


RK45 = RungeKutta45  # README.md:23-47 spelling


class Ensemble:
    """Batched builder: one instance per column of structure-of-arrays
    inputs.  Setters accept a scalar (broadcast) or one value per instance and
    validate every element like the single-instance setter."""

    def __init__(self, method: int, dim: int, n: int):
        self.method, self.dim_, self.n = method, dim, n
        h = C.c_void_p()
        _check(L.lib().bg_ensemble_new(method, dim, n, C.byref(h)))
        self._h = h
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().bg_ensemble_destroy(h)
            except Exception:
                pass
            self._h = None

    def _arr(self, v):
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))
        self._keep.append(a)
        return a

    def _set(self, fn, v):
        a = self._arr(v)
        _check(getattr(L.lib(), fn)(self._h, _dptr(a), a.size))
        return self

    def with_tolerance(self, v):
        return self._set("bg_ensemble_with_tolerance", v)

    def with_maximum_dt(self, v):
        return self._set("bg_ensemble_with_maximum_dt", v)

    def with_minimum_dt(self, v):
        return self._set("bg_ensemble_with_minimum_dt", v)

    def with_initial_time(self, v):
        return self._set("bg_ensemble_with_initial_time", v)

    def with_ending_time(self, v):
        return self._set("bg_ensemble_with_ending_time", v)

    def with_initial_conditions(self, y0_soa):
        a = self._arr(np.asarray(y0_soa, dtype=np.float64).reshape(-1))
        _check(L.lib().bg_ensemble_with_initial_conditions(self._h, _dptr(a), a.size))
        return self

    def with_derivative(self, rhs, params_soa=None, shared=None):
        """params_soa: [P, n] per-instance parameters (structure of arrays,
        P * n values), or None for a right-hand side without parameters."""
        if params_soa is not None:
            shape = np.shape(params_soa)
            if len(shape) != 2 or shape[1] != self.n:
                raise ValueError(f"params_soa must have shape (P, {self.n}); got {shape}")
        p = self._arr(np.zeros(0) if params_soa is None else np.asarray(params_soa).reshape(-1))
        n_params = 0 if params_soa is None else int(np.asarray(params_soa).shape[0])
        s = self._arr(np.zeros(0) if shared is None else np.asarray(shared).reshape(-1))
        _check(L.lib().bg_ensemble_with_derivative(self._h, _rhs_id(rhs), _dptr(p), n_params,
                                                   _dptr(s), s.size))
        return self

    def with_semantics(self, semantics: int):
        _check(L.lib().bg_ensemble_with_semantics(self._h, semantics))
        return self

    def with_max_attempts(self, n: int):
        _check(L.lib().bg_ensemble_with_max_attempts(self._h, n))
        return self

    def with_lu_precision(self, bits: int):
        """BDF: arithmetic of the Jacobian inverse (64, or 32 with FP64
        residuals and Broyden iteration)."""
        _check(L.lib().bg_ensemble_with_lu_precision(self._h, bits))
        return self

    def solve(self, plain: bool = False):
        fn = L.lib().bg_ensemble_solve_plain if plain else L.lib().bg_ensemble_solve
        _check(fn(self._h))
        return self.results()

    def solve_ivp(self):
        """README.md:45-47 chain per instance: Adams5 -> RK45 -> BDF6.
        Returns the results plus `method_used` per instance."""
        used = np.full(self.n, -1, dtype=np.int32)
        _check(L.lib().bg_ensemble_solve_ivp(self._h,
                                             used.ctypes.data_as(C.POINTER(C.c_int32))))
        out = self.results()
        out["method_used"] = used
        return out

    def results(self):
        t = np.zeros(self.n)
        y = np.zeros((self.dim_, self.n))
        st = np.zeros(self.n, dtype=np.int32)
        cnt = (L.Counts * max(self.n, 1))()
        _check(L.lib().bg_ensemble_results(self._h, _dptr(t), _dptr(y),
                                           st.ctypes.data_as(C.POINTER(C.c_int32)), cnt))
        c = np.ctypeslib.as_array(cnt).view(np.uint32).reshape(-1, 2)[: self.n]
        out = dict(t=t, y=y, status=st, n_accepted=c[:, 0].astype(np.int64),
                   n_rejected=c[:, 1].astype(np.int64))
        if self.method in (L.BDF6, L.BDF2):
            wk = (L.Counts * max(self.n, 1))()
            _check(L.lib().bg_ensemble_work(self._h, wk))
            w = np.ctypeslib.as_array(wk).view(np.uint32).reshape(-1, 2)[: self.n]
            out.update(n_newton=w[:, 0].astype(np.int64), n_lu=w[:, 1].astype(np.int64))
        return out


def solve_ensemble(method, w: dict, semantics=L.INTENDED, plain=False, max_attempts=0,
                   lu_bits=64):
    """Solve a workload dict (see bacon_amd.workloads) through the batched
    builder; returns final t, y, status and counters."""
    n = w["y0"].shape[1]
    e = (Ensemble(method, w["dim"], n)
         .with_minimum_dt(w["dt_min"]).with_maximum_dt(w["dt_max"])
         .with_tolerance(w["tol"])
         .with_initial_time(w.get("t0", 0.0)).with_ending_time(w["t_end"])
         .with_initial_conditions(w["y0"])
         .with_derivative(w["rhs"], w["params"] if w["params"].size else None,
                          w.get("shared"))
         .with_semantics(semantics))
    if max_attempts:
        e.with_max_attempts(max_attempts)
    if lu_bits != 64:
        e.with_lu_precision(lu_bits)
    return e.solve(plain=plain)


def solve_path(method, w: dict, capacity: int, every: int = 1, stops=None,
               stops_only: bool = False, semantics=L.INTENDED, device: str = "cuda:0",
               instance_major: bool = False):
    """Ensemble solve with path output (bg_rk_solve_path_device) for a
    workload dict (see bacon_amd.workloads).  Returns numpy arrays: final t,
    y, status, counts, and the per-instance ring buffers path_t[capacity][n],
    path_y[capacity][dim][n], path_len[n].  instance_major selects the
    [n][capacity] device layout (BG_PATH_INSTANCE_MAJOR); the arrays are
    returned in the same [capacity][...][n] shape either way."""
    import torch
    n, dim = w["y0"].shape[1], w["dim"]

    def T(a, dt=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)

    def arr(v):
        return np.array(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))

    keep = dict(y0=T(w["y0"]), params=T(w["params"] if w["params"].size else np.zeros((1, n))),
                t0=T(arr(w.get("t0", 0.0))), t_end=T(arr(w["t_end"])), tol=T(arr(w["tol"])),
                shared=T(w["shared"]) if w.get("shared") is not None else None,
                t=torch.zeros(n, dtype=torch.float64, device=device),
                y=torch.zeros((dim, n), dtype=torch.float64, device=device),
                st=torch.full((n,), -1, dtype=torch.int32, device=device),
                cnt=torch.zeros((n, 2), dtype=torch.int32, device=device),
                pt=torch.zeros((n, max(capacity, 1)) if instance_major else (max(capacity, 1), n),
                               dtype=torch.float64, device=device),
                py=torch.zeros((n, max(capacity, 1), dim) if instance_major
                               else (max(capacity, 1), dim, n), dtype=torch.float64, device=device),
                pl=torch.zeros(n, dtype=torch.int32, device=device))
    st_arr = None
    if stops is not None and len(stops):
        st_arr = T(np.asarray(stops, dtype=np.float64))
    d = L.EnsembleDesc(
        n=n, dim=dim, n_params=w["params"].shape[0], y0=keep["y0"].data_ptr(),
        params=keep["params"].data_ptr(),
        shared=keep["shared"].data_ptr() if keep["shared"] is not None else None,
        t0=keep["t0"].data_ptr(), t_end=keep["t_end"].data_ptr(), tol=keep["tol"].data_ptr(),
        dt_min_s=w["dt_min"], dt_max_s=w["dt_max"],
        t_out=keep["t"].data_ptr(), y_out=keep["y"].data_ptr(), status=keep["st"].data_ptr(),
        counts=keep["cnt"].data_ptr())
    p = L.PathDesc(capacity=capacity, every=every,
                   n_stops=0 if st_arr is None else st_arr.numel(),
                   flags=(L.PATH_STOPS_ONLY if stops_only else 0)
                   | (L.PATH_INSTANCE_MAJOR if instance_major else 0),
                   stops=None if st_arr is None else st_arr.data_ptr(),
                   path_t=keep["pt"].data_ptr(), path_y=keep["py"].data_ptr(),
                   path_len=keep["pl"].data_ptr())
    _check(L.lib().bg_rk_solve_path_device(method, _rhs_id(w["rhs"]), semantics, C.byref(d),
                                           C.byref(p), None))
    torch.cuda.synchronize()
    cnt = keep["cnt"].cpu().numpy().astype(np.int64)
    pt, py = keep["pt"].cpu().numpy(), keep["py"].cpu().numpy()
    if instance_major:
        pt, py = pt.T, py.transpose(1, 2, 0)
    return dict(t=keep["t"].cpu().numpy(), y=keep["y"].cpu().numpy(),
                status=keep["st"].cpu().numpy(), n_accepted=cnt[:, 0], n_rejected=cnt[:, 1],
                path_t=pt[:capacity], path_y=py[:capacity],
                path_len=keep["pl"].cpu().numpy().astype(np.int64))
