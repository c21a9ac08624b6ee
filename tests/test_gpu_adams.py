"""GPU Adams5 / Adams3 ensembles (bg_adams_solve_device, bg_ensemble_* with
BG_ADAMS5/3) against the oracle's Adams restatement, DEVICE profile: final
t, y, status and yield / rejection counts bit-identical, both semantics.
Right-hand sides that call cos are left out of the bit-exact set (the
device and host cos differ in the last bits)."""
import numpy as np
import pytest

import oracle as O
from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

SEM = [(L.AS_BUILT, O.AS_BUILT), (L.INTENDED, O.INTENDED)]


def _check(w, g, method_o, sem_o, idx, max_attempts=0):
    for i in idx:
        r = O.solve(method_o, w["rhs"], w["y0"][:, i], 0.0, float(np.broadcast_to(w["t_end"], (w["y0"].shape[1],))[i]),
                    float(np.broadcast_to(w["tol"], (w["y0"].shape[1],))[i]), w["dt_min"], w["dt_max"],
                    params=w["params"][:, i] if w["params"].size else None,
                    semantics=sem_o, profile=O.PROFILE_DEVICE, max_attempts=max_attempts)
        assert g["status"][i] == r["status"], (i, g["status"][i], r["status"])
        assert g["t"][i] == r["t"], i
        assert np.array_equal(g["y"][:, i], r["y"]), i
        assert g["n_accepted"][i] == r["n_accepted"] and g["n_rejected"][i] == r["n_rejected"], i


@pytest.mark.parametrize("sem,sem_o", SEM)
def test_adams5_lorenz_bit_exact(gpu, orc, sem, sem_o):
    """256 Lorenz-63 instances (t in [0,10], tol 1e-6).  As built, the stale
    derivative history drives most instances into MinimumTimeDeltaExceeded;
    the GPU reproduces those failures bit for bit as well."""
    w = W.lorenz_plain(256)
    w["dt_min"] = 1e-8
    g = ivp.solve_ensemble(L.ADAMS5, w, semantics=sem, max_attempts=200000)
    if sem == L.INTENDED:
        assert np.all(g["status"] == 0)
    _check(w, g, O.ADAMS5, sem_o, range(0, 256, 5), max_attempts=200000)


@pytest.mark.parametrize("sem,sem_o", SEM)
def test_adams3_lotka_volterra_bit_exact(gpu, orc, sem, sem_o):
    w = W.lotka_volterra(192)
    w["t_end"] = 15.0
    w["tol"] = 1e-7
    g = ivp.solve_ensemble(L.ADAMS3, w, semantics=sem, max_attempts=2000000)
    _check(w, g, O.ADAMS3, sem_o, range(0, 192, 7), max_attempts=2000000)


@pytest.mark.parametrize("method,method_o", [(L.ADAMS5, O.ADAMS5), (L.ADAMS3, O.ADAMS3)])
def test_adams_reference_problems(gpu, orc, method, method_o):
    """The reference's exp and quadratic Adams problems (adams.rs:696-922) as a
    mixed ensemble: bit-identical to the oracle, within the tests' epsilon of
    the closed form at t_end."""
    cases = [("exp", 1.0, 2.0, 1e-5, 0.1, 5e-4 if method == L.ADAMS5 else 1e-3, np.exp),
             ("quadratic", 1.0, 5.0, 1e-7 if method == L.ADAMS5 else 1e-5, 0.001,
              0.01 if method == L.ADAMS5 else 0.1, lambda t: 1.0 - t * t)]
    for rhs, y0, t_end, dt_min, dt_max, tol, exact in cases:
        n = 64
        w = dict(rhs=rhs, dim=1, y0=np.full((1, n), y0), params=np.zeros((0, n)), t0=0.0,
                 t_end=np.full(n, t_end), tol=np.full(n, tol), dt_min=dt_min, dt_max=dt_max)
        for sem, sem_o in SEM:
            g = ivp.solve_ensemble(method, w, semantics=sem)
            assert np.all(g["status"] == 0)
            assert np.all(np.abs(g["y"][0] - exact(t_end)) <= 0.01)
            _check(w, g, method_o, sem_o, [0, 63])


def _oracle_path(method_o, rhs, y0, t_end, tol, dt_min, dt_max, sem_o, params=None, cap=400000):
    return O.solve(method_o, rhs, y0, 0.0, t_end, tol, dt_min, dt_max, params=params,
                   semantics=sem_o, profile=O.PROFILE_DEVICE, path_cap=cap)


def _assert_path(path, r):
    assert len(path) == r["path_len"]
    assert np.array_equal(np.array([p[0] for p in path]), r["path_t"])
    assert np.array_equal(np.array([p[1] for p in path]).reshape(r["path_y"].shape), r["path_y"])


@pytest.mark.parametrize("sem,sem_o", SEM)
@pytest.mark.parametrize("cls", [ivp.Adams5, ivp.Adams3])
def test_single_instance_reference_problems(gpu, orc, cls, sem, sem_o):
    """adams.rs:713-922 through the single-instance builder and iterator:
    every yield (the bootstrap replays of adams.rs:415-428 included) equals
    the oracle's path bit for bit and stays on the exact solution."""
    method_o = O.ADAMS5 if cls is ivp.Adams5 else O.ADAMS3
    a5 = cls is ivp.Adams5
    cases = [("exp", 1.0, 2.0, 1e-5, 0.1, 5e-4 if a5 else 1e-3, np.exp, 0.01),
             ("quadratic", 1.0, 5.0, 1e-7 if a5 else 1e-5, 0.001, 0.01 if a5 else 0.1,
              lambda t: 1.0 - t * t, 0.01)]
    for rhs, y0, t_end, dt_min, dt_max, tol, exact, bar in cases:
        s = (cls.new(1).with_minimum_dt(dt_min).with_maximum_dt(dt_max).with_tolerance(tol)
             .with_initial_time(0.0).with_ending_time(t_end)
             .with_initial_conditions_slice([y0]).with_derivative(rhs).with_semantics(sem))
        path = s.solve().collect_vec()
        assert len(path) > 0
        for t, y in path:
            assert abs(y[0] - exact(t)) <= bar * max(1.0, abs(exact(t)))
        _assert_path(path, _oracle_path(method_o, rhs, [y0], t_end, tol, dt_min, dt_max, sem_o))


def test_single_instance_lorenz_and_as_built_failure(gpu, orc):
    """Adams5 on one Lorenz-63 instance: intended runs to the end; as built
    the stale derivative history ends the path with MinimumTimeDeltaExceeded,
    after the same yields as the oracle's, and the iterator is fused."""
    w = W.lorenz_plain(1)
    y0, p = w["y0"][:, 0], w["params"][:, 0]

    def mk(sem):
        return (ivp.Adams5.new(3).with_minimum_dt(1e-6).with_maximum_dt(0.1)
                .with_tolerance(1e-6).with_initial_time(0.0).with_ending_time(10.0)
                .with_initial_conditions(y0).with_derivative("lorenz", p).with_semantics(sem))
    path = mk(L.INTENDED).solve().collect_vec()
    r = _oracle_path(O.ADAMS5, "lorenz", y0, 10.0, 1e-6, 1e-6, 0.1, O.INTENDED, params=p)
    assert r["status"] == 0 and path[-1][0] == 10.0
    _assert_path(path, r)
    it = mk(L.AS_BUILT).solve()
    vals = []
    with pytest.raises(ivp.IVPError) as e:
        for v in it:
            vals.append(v)
    assert e.value.name == "MinimumTimeDeltaExceeded"
    assert list(it) == []
    r = _oracle_path(O.ADAMS5, "lorenz", y0, 10.0, 1e-6, 1e-6, 0.1, O.AS_BUILT, params=p)
    assert r["status"] == O.MINIMUM_TIME_DELTA_EXCEEDED
    _assert_path(vals, r)


def test_single_instance_path_resumes_across_buffers(gpu, orc):
    """More yields than one launch's path buffer (65,536): the lane state is
    saved and the kernel relaunched; the joined path equals the oracle's."""
    s = (ivp.Adams3.new(1).with_minimum_dt(1e-6).with_maximum_dt(1e-3).with_tolerance(1e-3)
         .with_initial_time(0.0).with_ending_time(100.0)
         .with_initial_conditions_slice([1.0]).with_derivative("neg"))
    path = s.solve().collect_vec()
    assert len(path) > 70000
    _assert_path(path, _oracle_path(O.ADAMS3, "neg", [1.0], 100.0, 1e-3, 1e-6, 1e-3,
                                    O.INTENDED))


def test_single_instance_missing_parameters(gpu):
    s = ivp.Adams5.new(1).with_tolerance(1e-4).with_minimum_dt(1e-5)
    with pytest.raises(ivp.IVPError) as e:
        s.solve()
    assert e.value.name == "MissingParameters"
