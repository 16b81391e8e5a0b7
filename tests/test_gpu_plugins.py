"""GPU parity of the remaining plug-ins: continuous functions, job scheduling,
ZDT1, multi-day VRP and bi-objective VRP, called through the C ABI and
compared with the oracle (which test_oracle_plugins.py pins to the
reference's own VRP code and to FP64 formulas).

Bar: decoded rows bit-exact; fitness / objectives bit-exact to the oracle
(same FP32 terms, order-independent fixed-point sums, same deterministic
cosine); evolution trajectories (keys, fitness, rank orders, NSGA-II ranks
and crowding) bit-exact.
"""
import numpy as np
import pytest

import brkga_amd
import pyoracle
from cases import make_problem
from conftest import random_keys
from test_gpu_parity import compare_brkga, compare_nsga, gpu_evaluate

pytestmark = pytest.mark.gpu

EVAL_CASES = [(k, n) for k in ("cont_sphere", "cont_rastrigin", "cont_rosenbrock") for n in (1, 2, 30, 1000)] + [
    ("jobshop", 1), ("jobshop", 50), ("jobshop", 2000),
    ("zdt1", 1), ("zdt1", 2), ("zdt1", 30), ("zdt1", 700),
    ("vrp_days", 1), ("vrp_days", 25), ("vrp_days", 300), ("vrp_days_tight", 60),
    ("movrp", 1), ("movrp", 31), ("movrp", 200)]


@pytest.mark.parametrize("kind,n", EVAL_CASES)
def test_plugin_evaluate_and_decode_parity(engine, kind, n):
    print(f"plugin evaluate parity {kind} n={n}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=n)
    L = op.chrom_len
    for ties in (False, True):
        keys = random_keys(n * 5 + ties, 40, L, ties=ties)
        gf, gd = gpu_evaluate(gp, keys)
        of, od = op.evaluate(keys)
        assert np.array_equal(gd, od), f"decode mismatch ({kind}, n={n}, ties={ties})"
        assert np.array_equal(gf, of), f"fitness mismatch ({kind}, n={n}, ties={ties})"


def test_job_scheduling_machine_limits(engine):
    t = np.ones(10, np.float32)
    for bad in (0, 4097):
        with pytest.raises(brkga_amd.BrkgaError) as e:
            brkga_amd.Problem.job_scheduling(engine, t, bad)
        assert e.value.status == brkga_amd.ERR_INVALID_ARG
    gp = brkga_amd.Problem.job_scheduling(engine, t, 4096)
    op = pyoracle.Problem.job_scheduling(t, 4096)
    keys = random_keys(5, 8, 10)
    gf, gd = gpu_evaluate(gp, keys)
    of, od = op.evaluate(keys)
    assert np.array_equal(gd, od) and np.array_equal(gf, of)


def test_plugin_rejects_bad_data(engine):
    with pytest.raises(brkga_amd.BrkgaError):
        brkga_amd.Problem.continuous(engine, 3, 10, -1.0, 1.0)
    with pytest.raises(brkga_amd.BrkgaError):
        brkga_amd.Problem.vrp_days(engine, np.zeros((1, 1), np.float32), 5, 100.0)
    with pytest.raises(brkga_amd.BrkgaError):
        brkga_amd.Problem.movrp(engine, np.zeros((4, 2), np.float32), np.array([0, 1, -2, 3], np.int32), 10, 2)


BRKGA_CASES = [
    # kind, n, P, E, M, rho, islands, interval, k, generations
    ("cont_sphere", 30, 64, 12, 6, 0.7, 1, 0, 0, 8),
    ("cont_rastrigin", 30, 4 * 32, 4 * 6, 4 * 5, 0.7, 4, 3, 2, 8),
    ("cont_rosenbrock", 20, 64, 12, 6, 0.7, 1, 0, 0, 8),
    ("jobshop", 50, 64, 12, 6, 0.7, 2, 2, 2, 8),
    ("vrp_days", 40, 64, 12, 6, 0.7, 1, 0, 0, 8),
    ("vrp_days_tight", 30, 48, 9, 5, 0.7, 1, 0, 0, 6),
]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("case", BRKGA_CASES, ids=lambda c: f"{c[0]}-n{c[1]}-I{c[6]}")
def test_plugin_brkga_trajectory_parity(engine, case, graph):
    kind, n, P, E, M, rho, I, iv, k, G = case
    print(f"plugin brkga trajectory {case} graph={graph}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=3)
    seed = 0xFACE + n
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E, M, rho, I, iv, k, seed))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, rho, I, iv, k, seed, op.maximize))
    compare_brkga(pop, ost, "after init")
    if graph:
        pop.evolve(G, graph=True)
        for _ in range(G):
            ost.step()
        compare_brkga(pop, ost, f"after {G} graph steps")
    else:
        for t in range(G):
            pop.evolve(1)
            ost.step()
            compare_brkga(pop, ost, f"after step {t}")


@pytest.mark.parametrize("kind,n,N,G", [("zdt1", 30, 100, 12), ("zdt1", 5, 32, 6), ("movrp", 31, 64, 8),
                                        ("movrp", 12, 32, 5)])
def test_plugin_nsga2_trajectory_parity(engine, kind, n, N, G):
    print(f"plugin nsga2 trajectory {kind} n={n} N={N}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=4)
    seed = 4242 + n
    pop = brkga_amd.Population(gp, brkga_amd.Params(N, 0, 0, 0.5, 1, 0, 0, seed))
    pop.initialize()
    ost = pyoracle.Nsga2State(op, pyoracle.params(N, 0, 0, 0.5, seed=seed))
    compare_nsga(pop, ost, "init")
    for t in range(G):
        pop.evolve(1)
        ost.step()
        compare_nsga(pop, ost, f"step {t}")


def test_zdt1_approaches_front(engine):
    """End-to-end quality: NSGA-II on ZDT1 (n = 30, N = 100, the reference
    example's size) moves its first front towards f2 = 1 - sqrt(f1)."""
    gp, _ = make_problem(engine, "zdt1", 30)
    pop = brkga_amd.Population(gp, brkga_amd.Params(100, 0, 0, 0.5, 1, 0, 0, 7))
    pop.initialize()
    r0 = pop.read()
    gap0 = np.mean(r0["fitness"][:, 1] - (1 - np.sqrt(r0["fitness"][:, 0])))
    pop.evolve(150)
    r = pop.read()
    front = r["rank"] == 0
    f = r["fitness"][front]
    gap = np.mean(f[:, 1] - (1 - np.sqrt(f[:, 0])))
    assert gap > -1e-6 and gap < 0.5 * gap0, f"front gap {gap} vs initial {gap0}"
