"""The oracle's restatements of the remaining plug-ins, pinned.

- VRP days and the bi-objective VRP: against the reference's own
  VrpBlazerConfig / MultiObjectiveVRPConfig compiled into oracle/_ref
  (configs/vrp_blazer_config.hpp:192-296,
  configs/multi_objective_vrp_config.hpp:69-300): identical day / route
  assignment, objectives within FP32 summation error.
- Continuous functions, job scheduling, ZDT1: their reference headers need
  CUDA, so they are pinned by FP64 recomputation of the formulas
  (configs/continuous_opt_gpu_config.hpp:132-160,
  configs/job_scheduling_gpu_config.hpp:138-158, examples/zdt1.cu:43-63)
  and by closed forms at known optima.
- The deterministic cosine: against numpy's FP64 cosine.
"""
import os
import tempfile

import numpy as np
import pytest

import pyoracle

RNG = np.random.default_rng(20261204)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def test_cos_det_matches_fp64_cosine():
    y = np.concatenate([RNG.uniform(-40, 40, 20000), np.arange(-64, 65) * (np.pi / 4),
                        [0.0, 1e-8, -1e-8, 6.2831855, 32.16991]]).astype(np.float32)
    got = np.array([pyoracle.cos_det(v) for v in y], np.float32)
    want = np.cos(y.astype(np.float64))
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - want) / ulp
    assert err.max() <= 0.5 + 1e-3, f"worst error {err.max():.3f} ulp at {y[err.argmax()]}"
    assert pyoracle.cos_det(0.0) == 1.0


def x_of(keys, lo, hi):
    w = np.float32(hi) - np.float32(lo)
    return (keys.astype(np.float32) * w).astype(np.float32) + np.float32(lo)


@pytest.mark.parametrize("fn", [pyoracle.SPHERE, pyoracle.RASTRIGIN, pyoracle.ROSENBROCK])
def test_continuous_against_fp64(fn):
    lo, hi = (-5.0, 10.0) if fn == pyoracle.ROSENBROCK else (-5.12, 5.12)
    n = 37
    p = pyoracle.Problem.continuous(fn, n, lo, hi)
    keys = RNG.random((50, n), dtype=np.float32)
    fit, dec = p.evaluate(keys)
    x = x_of(keys, lo, hi).astype(np.float64)
    assert np.array_equal(dec.view(np.float32), x_of(keys, lo, hi))
    if fn == pyoracle.SPHERE:
        want = (x * x).sum(1)
        scale = want
    elif fn == pyoracle.RASTRIGIN:
        terms = x * x - 10 * np.cos(2 * np.pi * x)
        want = 10 * n + terms.sum(1)
        scale = 10 * n + np.abs(terms).sum(1)
    else:
        t = 100 * (x[:, 1:] - x[:, :-1] ** 2) ** 2 + (1 - x[:, :-1]) ** 2
        want = t.sum(1)
        scale = want
    assert (np.abs(fit - want) / scale).max() < 2e-6


def test_continuous_optima():
    n = 20
    half = np.full((1, n), 0.5, np.float32)
    for fn in (pyoracle.SPHERE, pyoracle.RASTRIGIN):
        fit, _ = pyoracle.Problem.continuous(fn, n, -5.12, 5.12).evaluate(half)
        assert fit[0] == 0.0  # x = 0 exactly
    fit, _ = pyoracle.Problem.continuous(pyoracle.ROSENBROCK, n, 0.0, 2.0).evaluate(half)
    assert fit[0] == 0.0  # x = 1 exactly
    # a single variable: Rosenbrock has no terms
    fit, _ = pyoracle.Problem.continuous(pyoracle.ROSENBROCK, 1, -5.0, 10.0).evaluate(np.array([[0.3]], np.float32))
    assert fit[0] == 0.0


def test_job_scheduling_against_fp64():
    n, M = 61, 7
    times = RNG.uniform(1, 50, n).astype(np.float32)
    p = pyoracle.Problem.job_scheduling(times, M)
    keys = RNG.random((40, n), dtype=np.float32)
    keys[0, :] = 0.9999999  # everything on the last machine
    fit, dec = p.evaluate(keys)
    mach = np.minimum((keys * np.float32(M)).astype(np.int64), M - 1)
    assert np.array_equal(dec, mach)
    for r in range(keys.shape[0]):
        loads = np.bincount(mach[r], weights=times.astype(np.float64), minlength=M)
        assert rel(fit[r], loads.max()) < 1e-6
    assert rel(fit[0], times.astype(np.float64).sum()) < 1e-6


def test_job_scheduling_integer_times_exact():
    times = np.arange(10, 50, dtype=np.float32)  # create_test_instance pattern
    p = pyoracle.Problem.job_scheduling(times, 5)
    keys = RNG.random((20, times.size), dtype=np.float32)
    fit, dec = p.evaluate(keys)
    for r in range(20):
        assert fit[r] == np.bincount(dec[r], weights=times, minlength=5).max()


def test_zdt1_front_and_formula():
    n = 30
    p = pyoracle.Problem.zdt1(n)
    keys = RNG.random((64, n), dtype=np.float32)
    keys[:8, 1:] = 0.0  # g = 1: on the true front f2 = 1 - sqrt(f1)
    obj, dec = p.evaluate(keys)
    assert np.array_equal(dec.view(np.float32), keys)
    x = keys.astype(np.float64)
    g = 1 + 9 * x[:, 1:].sum(1) / (n - 1)
    f2 = g * (1 - np.sqrt(x[:, 0] / g))
    assert np.array_equal(obj[:, 0], keys[:, 0])
    assert rel(obj[:, 1], f2).max() < 1e-6
    assert np.allclose(obj[:8, 1], 1 - np.sqrt(x[:8, 0]), rtol=1e-6, atol=0)


def vrp_instance(n, seed):
    r = np.random.default_rng(seed)
    xy = r.uniform(0, 100, (n + 1, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    return (d * r.uniform(0.8, 1.2, d.shape)).astype(np.float32)


@pytest.mark.parametrize("n,stops,hours,seed", [(8, 3, 4.0, 1), (25, 12, 8.0, 2), (40, 5, 3.0, 3), (30, 0, 8.0, 4),
                                                (20, 4, 1.2, 5)])
def test_vrp_days_against_reference(n, stops, hours, seed):
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    tm = vrp_instance(n, seed)
    keys = RNG.random((30, n), dtype=np.float32)
    keys[0, : n // 2] = keys[0, 0]  # ties broken by stop index
    p = pyoracle.Problem.vrp_days(tm, stops, hours * 60.0)
    fit, dec = p.evaluate(keys)
    rfit = np.empty(30, np.float32)
    rday = np.empty((30, n), np.int32)
    ref.ref_vrp_days(pyoracle._f(tm), n + 1, stops, hours, pyoracle._f(keys), 30, pyoracle._f(rfit),
                     pyoracle._i(rday))
    assert np.array_equal(dec, rday)
    assert rel(fit, rfit).max() < 5e-6
    if stops == 0:
        assert np.all(fit == 2000.0 * n)


def write_cvrp(path, xy, dem, cap, vehicles):
    dim = xy.shape[0]
    with open(path, "w") as f:
        f.write(f"NAME : T-n{dim}-k{vehicles}\nTYPE : CVRP\nDIMENSION : {dim}\nCAPACITY : {cap}\n")
        f.write("NODE_COORD_SECTION\n")
        for i in range(dim):
            f.write(f"{i + 1} {xy[i, 0]:.9g} {xy[i, 1]:.9g}\n")
        f.write("DEMAND_SECTION\n")
        for i in range(dim):
            f.write(f"{i + 1} {dem[i]}\n")
        f.write("DEPOT_SECTION\n1\n-1\nEOF\n")


@pytest.mark.parametrize("n,cap,vehicles,seed", [(31, 100, 5, 1), (12, 30, 2, 2), (50, 1000, 3, 3), (20, 15, 9, 4)])
def test_movrp_against_reference(n, cap, vehicles, seed):
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    r = np.random.default_rng(seed)
    xy = np.round(r.uniform(0, 100, (n + 1, 2)), 3).astype(np.float32)
    dem = r.integers(1, 30, n + 1).astype(np.int32)
    dem[0] = 0
    dem[3] = cap + 7  # one customer heavier than a vehicle: overload penalty
    keys = RNG.random((30, n), dtype=np.float32)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "t.vrp").encode()
        write_cvrp(path.decode(), xy, dem, cap, vehicles)
        import ctypes
        v, c = ctypes.c_int(), ctypes.c_int()
        assert ref.ref_movrp_info(path, ctypes.byref(v), ctypes.byref(c)) == n
        assert (v.value, c.value) == (vehicles, cap)
        robj = np.empty((30, 2), np.float32)
        rroute = np.empty((30, n), np.int32)
        ref.ref_movrp(path, pyoracle._f(keys), 30, pyoracle._f(robj), pyoracle._i(rroute))
    p = pyoracle.Problem.movrp(xy, dem, cap, vehicles)
    obj, dec = p.evaluate(keys)
    assert np.array_equal(dec, rroute)
    assert rel(obj, robj).max() < 5e-6
