"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/bacon_gpu.h, and the host-side builder validation matches
the reference's setters (src/ivp/rk.rs:168-256) and the oracle's restatement.
No compute entry point is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bg_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    return L.lib()


def test_library_exports_every_declared_symbol(lib):
    names = _declared("bacon_gpu.h")
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers the header exactly
    assert sorted(L.exported_names()) == names


def test_builder_static_dynamic(lib):
    h = C.c_void_p()
    assert lib.bg_new(L.RK45, L.DIM_DYN, 1, C.byref(h)) == 12  # StaticOnDynamic
    assert lib.bg_new_dyn(L.RK45, L.DIM_STATIC, 3, C.byref(h)) == 11  # DynamicOnStatic
    assert lib.bg_new(L.RK45, L.DIM_STATIC, 3, C.byref(h)) == 0
    d = C.c_uint32()
    assert lib.bg_dim(h, C.byref(d)) == 0 and d.value == 3
    lib.bg_destroy(h)


def test_builder_validation_matches_oracle(lib, orc):
    """Random setter sequences give the same status codes through the C ABI
    and the oracle's builder restatement."""
    rng = np.random.default_rng(3)
    names = ["with_tolerance", "with_maximum_dt", "with_minimum_dt",
             "with_initial_time", "with_ending_time"]
    for trial in range(300):
        h = C.c_void_p()
        assert lib.bg_new(L.RK45, L.DIM_STATIC, 1, C.byref(h)) == 0
        ob = orc.OracleBuilder()
        assert ob.call("new", orc.RK45, 1) == 0
        for _ in range(rng.integers(1, 9)):
            k = rng.integers(0, len(names))
            v = float(rng.choice([-1.0, 0.0, 1e-3, 0.05, 0.1, 1.0, 2.0, 10.0]))
            got = getattr(lib, "bg_" + names[k])(h, v)
            exp = ob.call(names[k], v)
            assert got == exp, (trial, names[k], v)
        y = np.array([1.0])
        if rng.random() < 0.8:
            assert lib.bg_with_initial_conditions_slice(
                h, y.ctypes.data_as(C.POINTER(C.c_double)), 1) == 0
            ob.call("with_initial_conditions", 1)
        if rng.random() < 0.8:
            assert lib.bg_with_derivative(h, L.RHS["exp"], None, 0) == 0
            ob.call("with_derivative")
        expect = ob.call("solve")
        if expect == orc.MISSING_PARAMETERS:
            assert lib.bg_solve(h) == 1
        lib.bg_destroy(h)


def test_builder_python_mirror_raises(lib):
    s = ivp.RungeKutta45.new(1)
    with pytest.raises(ivp.IVPError) as e:
        s.with_tolerance(0.0)
    assert e.value.name == "ToleranceOOB"
    with pytest.raises(ivp.IVPError) as e:
        s.with_minimum_dt(-1)
    assert e.value.name == "TimeDeltaOOB"
    s.with_ending_time(1.0)
    with pytest.raises(ivp.IVPError) as e:
        s.with_initial_time(2.0)
    assert e.value.name == "TimeStartOOB"
    with pytest.raises(ivp.IVPError) as e:
        s.solve()
    assert e.value.name == "MissingParameters"
    with pytest.raises(ivp.IVPError) as e:
        s.with_initial_conditions_slice([1.0, 2.0])  # reference: panic
    assert e.value.name == "InvalidArgument"
    with pytest.raises(ivp.IVPError):
        ivp.RungeKutta45.new(1, dim_kind=L.DIM_DYN)


def test_readme_aliases(lib):
    s = ivp.RK45.new(1).with_dt_min(0.01).with_dt_max(0.1).with_tolerance(1e-4)
    s.with_initial_conditions([1.0]).with_start(0.0).with_end(10.0)
    with pytest.raises(ivp.IVPError) as e:
        s.with_end(-1.0)
    assert e.value.name == "TimeEndOOB"


def test_ensemble_setters_validate_each_element(lib):
    e = ivp.Ensemble(L.RK45, 3, 4)
    with pytest.raises(ivp.IVPError) as ex:
        e.with_tolerance([1e-6, 0.0, 1e-6, 1e-6])
    assert ex.value.name == "ToleranceOOB"
    with pytest.raises(ivp.IVPError) as ex:
        e.with_tolerance([1e-6, 1e-6])  # neither 1 nor n values
    assert ex.value.name == "InvalidArgument"
    e.with_initial_time(0.0)
    with pytest.raises(ivp.IVPError) as ex:
        e.with_ending_time([1.0, 2.0, 0.0, 3.0])
    assert ex.value.name == "TimeEndOOB"
    with pytest.raises(ivp.IVPError) as ex:
        e.solve()
    assert ex.value.name == "MissingParameters"


def test_workload_generator_is_deterministic():
    a = W.lorenz_batched(1000)
    b = W.lorenz_batched(1000)
    assert np.array_equal(a["y0"], b["y0"]) and np.array_equal(a["tol"], b["tol"])
    # an instance's draws do not depend on the batch it is generated in
    c = W.lorenz_batched(10, first=500)
    assert np.array_equal(c["y0"], a["y0"][:, 500:510])
    assert a["tol"].min() >= 1e-10 and a["tol"].max() <= 1e-4
    assert a["t_end"].min() >= 1.0 and a["t_end"].max() <= 20.0
    assert np.all((a["params"][1] >= 20) & (a["params"][1] <= 40))
    p = W.lorenz_plain(1024)
    assert p["y0"].shape == (3, 1024)
    assert np.all(p["y0"][2] >= 5) and np.all(p["y0"][2] <= 40)


def test_cost_order_longest_first():
    t_end = np.array([1.0, 20.0, 5.0])
    tol = np.array([1e-4, 1e-4, 1e-10])
    o = W.cost_order(t_end, tol)
    assert list(o) == [2, 1, 0]


def test_ensemble_linear_shared_block_size(lib):
    """The linear right-hand side needs the whole [A; b] block (dim^2 + dim)."""
    e = ivp.Ensemble(L.RK45, 128, 2)
    with pytest.raises(ivp.IVPError) as ex:
        e.with_derivative("linear", shared=np.zeros(128 * 128))
    assert ex.value.name == "InvalidArgument"
    e.with_derivative("linear", shared=np.zeros(128 * 128 + 128))
