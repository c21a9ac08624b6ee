"""ctypes binding of libbacon_gpu.so (the C ABI declared in include/bacon_gpu.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BACON_GPU_LIB may name another build of the same sources (the host-side
# sanitizer build, libbacon_gpu_san.so); there is no other implementation.
LIB_PATH = os.environ.get("BACON_GPU_LIB") or os.path.join(HERE, "libbacon_gpu.so")

OK = 0
MINIMUM_TIME_DELTA_EXCEEDED = 8
MAXIMUM_ITERATIONS_EXCEEDED = 9
SINGULAR_MATRIX = 10
DONE = 100
ERR_INVALID_ARGUMENT = -1
ERR_HIP = -2
ERR_UNSUPPORTED = -3
ERR_STATE = -4

RK45, RK23, BDF6, BDF2, ADAMS5, ADAMS3 = 0, 1, 2, 3, 4, 5
AS_BUILT, INTENDED = 0, 1
DIM_STATIC, DIM_DYN = 0, 1

RHS = {
    "exp": 0, "neg": 1, "quadratic": 2, "cos": 3, "lorenz": 4, "linear": 5,
    "brusselator": 6, "lotka_volterra": 7, "oscillator": 8, "fail_after": 9, "robertson": 10, "two_level": 11,
}

_dp = C.POINTER(C.c_double)


class Counts(C.Structure):
    _fields_ = [("n_accepted", C.c_uint32), ("n_rejected", C.c_uint32)]


class Launch(C.Structure):
    _fields_ = [("block", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("refill_min", C.c_uint32), ("prio_cutoff", C.c_uint32)]


class EnsembleDesc(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("dim", C.c_uint32), ("n_params", C.c_uint32),
        ("y0", C.c_void_p), ("params", C.c_void_p), ("shared", C.c_void_p),
        ("t0", C.c_void_p), ("t_end", C.c_void_p), ("tol", C.c_void_p),
        ("dt_min", C.c_void_p), ("dt_max", C.c_void_p),
        ("t0_s", C.c_double), ("t_end_s", C.c_double), ("tol_s", C.c_double),
        ("dt_min_s", C.c_double), ("dt_max_s", C.c_double),
        ("order", C.c_void_p), ("max_attempts", C.c_uint64),
        ("t_out", C.c_void_p), ("y_out", C.c_void_p), ("status", C.c_void_p),
        ("counts", C.c_void_p), ("work", C.c_void_p),
    ]


PATH_STOPS_ONLY = 1
PATH_INSTANCE_MAJOR = 2


class PathDesc(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32), ("every", C.c_uint32), ("n_stops", C.c_uint32),
        ("flags", C.c_uint32), ("stops", C.c_void_p), ("path_t", C.c_void_p),
        ("path_y", C.c_void_p), ("path_len", C.c_void_p),
    ]


TUNE_MAX_PARAMS = 8


class Tuning(C.Structure):
    _fields_ = [
        ("launch", Launch), ("n_params", C.c_uint32), ("pilot_n", C.c_uint32),
        ("coef", C.c_double * (3 + TUNE_MAX_PARAMS)),
        ("pilot_ms", C.c_double), ("default_ms", C.c_double), ("best_ms", C.c_double),
        ("n_trials", C.c_uint32), ("pad", C.c_uint32),
    ]

    def as_dict(self):
        return {"block": self.launch.block, "blocks_per_cu": self.launch.blocks_per_cu,
                "coef": [float(c) for c in self.coef[: 3 + self.n_params]],
                "pilot_n": self.pilot_n, "pilot_ms": self.pilot_ms,
                "default_ms": self.default_ms, "best_ms": self.best_ms,
                "n_trials": self.n_trials}


class AdjointDesc(C.Structure):
    _fields_ = [
        ("base", EnsembleDesc), ("n_obs", C.c_uint32), ("checkpoint_every", C.c_uint32),
        ("max_checkpoints", C.c_uint32), ("blocks_per_cu", C.c_uint32),
        ("obs_t", C.c_void_p), ("targets", C.c_void_p), ("loss", C.c_void_p),
        ("grad_params", C.c_void_p), ("grad_y0", C.c_void_p), ("obs_y", C.c_void_p),
    ]


class AdjointTuning(C.Structure):
    _fields_ = [("checkpoint_every", C.c_uint32), ("max_checkpoints", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("n_trials", C.c_uint32),
                ("default_ms", C.c_double), ("best_ms", C.c_double),
                ("schedule", C.c_uint32), ("pad", C.c_uint32), ("forward", Tuning)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "forward")}
        if self.schedule:
            f = self.forward
            d["forward_model"] = {"coef": [float(c) for c in f.coef[:3 + f.n_params]],
                                  "pilot_n": f.pilot_n}
        return d


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("bg_new", C.c_int, [C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("bg_new_dyn", C.c_int, [C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("bg_dim", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("bg_with_tolerance", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_maximum_dt", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_minimum_dt", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_initial_time", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_ending_time", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_initial_conditions_slice", C.c_int, [C.c_void_p, _dp, C.c_size_t]),
    ("bg_with_derivative", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_size_t]),
    ("bg_with_semantics", C.c_int, [C.c_void_p, C.c_int]),
    ("bg_with_dt_min", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_dt_max", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_start", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_with_end", C.c_int, [C.c_void_p, C.c_double]),
    ("bg_solve", C.c_int, [C.c_void_p]),
    ("bg_next", C.c_int, [C.c_void_p, _dp, _dp]),
    ("bg_time", C.c_int, [C.c_void_p, _dp]),
    ("bg_destroy", None, [C.c_void_p]),
    ("bg_ensemble_new", C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("bg_ensemble_with_tolerance", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_maximum_dt", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_minimum_dt", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_initial_time", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_ending_time", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_initial_conditions", C.c_int, [C.c_void_p, _dp, C.c_uint64]),
    ("bg_ensemble_with_derivative", C.c_int,
     [C.c_void_p, C.c_int, _dp, C.c_uint32, _dp, C.c_uint64]),
    ("bg_ensemble_with_semantics", C.c_int, [C.c_void_p, C.c_int]),
    ("bg_ensemble_with_max_attempts", C.c_int, [C.c_void_p, C.c_uint64]),
    ("bg_ensemble_with_lu_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("bg_ensemble_solve", C.c_int, [C.c_void_p]),
    ("bg_ensemble_solve_ivp", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("bg_ensemble_solve_plain", C.c_int, [C.c_void_p]),
    ("bg_ensemble_results", C.c_int,
     [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32), C.POINTER(Counts)]),
    ("bg_ensemble_work", C.c_int, [C.c_void_p, C.POINTER(Counts)]),
    ("bg_ensemble_destroy", None, [C.c_void_p]),
    ("bg_rk_solve_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Launch), C.c_void_p]),
    ("bg_rk_solve_device_noalloc", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Launch), C.c_void_p]),
    ("bg_rk_solve_path_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(PathDesc), C.c_void_p]),
    ("bg_rk_autotune_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Tuning), C.c_void_p]),
    ("bg_rk_schedule_device", C.c_int,
     [C.POINTER(Tuning), C.POINTER(EnsembleDesc), C.c_void_p, C.c_void_p]),
    ("bg_rk_solve_tuned_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Tuning), C.c_void_p]),
    ("bg_schedule_by_counts_device", C.c_int,
     [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    ("bg_graph_capture_rk", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Launch),
      C.POINTER(C.c_void_p)]),
    ("bg_graph_capture_bdf", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(C.c_void_p)]),
    ("bg_graph_launch", C.c_int, [C.c_void_p, C.c_void_p]),
    ("bg_graph_destroy", None, [C.c_void_p]),
    ("bg_rk_solve_plain_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.c_void_p]),
    ("bg_bdf_solve_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.c_void_p]),
    ("bg_bdf_autotune_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Tuning), C.c_void_p]),
    ("bg_bdf_solve_tuned_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.POINTER(Tuning), C.c_void_p]),
    ("bg_bdf_inverse_device", C.c_int,
     [C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("bg_adams_solve_device", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.POINTER(EnsembleDesc), C.c_void_p]),
    ("bg_rk_vjp_device", C.c_int,
     [C.c_int, C.c_int, C.POINTER(AdjointDesc), C.c_void_p]),
    ("bg_adjoint_solve_tuned_device", C.c_int,
     [C.c_int, C.c_int, C.POINTER(AdjointDesc), C.POINTER(AdjointTuning), C.c_void_p]),
    ("bg_adjoint_autotune_device", C.c_int,
     [C.c_int, C.c_int, C.POINTER(AdjointDesc), C.POINTER(AdjointTuning), C.c_void_p]),
    ("bg_version", C.c_int, []),
    ("bg_device_info", C.c_int,
     [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load libbacon_gpu.so; raise if it is absent (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m bacon_amd.build` "
                "(the GPU path has no fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_names():
    return [s[0] for s in SIGNATURES]
