"""bacon_amd — MI355X-native batched adaptive IVP engine.

The hot path of bacon-sci's `ivp` module (RK45 stage loop / error estimate /
step controller, BDF6 bootstrap / Broyden corrector / controller) as
hand-written gfx950 HIP kernels behind a C ABI (include/bacon_gpu.h), with a
host-side mirror of the reference's builder API in `bacon_amd.ivp`.
"""
from . import _lib  # noqa: F401
from .ivp import (BDF2, BDF6, RK45, Ensemble, IVPError, IVPIterator,  # noqa: F401
                  RungeKutta23, RungeKutta45, solve_ensemble)
