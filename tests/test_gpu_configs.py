"""GPU parity at the benchmark configurations, each at its own workload.

  * C1  TSP n = 52 (coordinates U[0,1000)^2), P = 256, E = 51, M = 25,
        rho 0.7, 500 generations: the oracle, every 50 generations.
  * C3a greedy knapsack n = 500, P = 2048 as ONE population (wider than any
        workgroup, so every generation runs the partition launch, with
        maximisation), E = 409, M = 204, rho 0.7: the oracle, 30 generations.
  * C3b TSPJ n = 200 (chromosome 400), P = 1024, E = 204, M = 102,
        rho 0.75: the oracle, 20 generations.
  * C4  the 512-instance batched set (n log-uniform in [50, 2000], P = 256,
        E = M = 38): the captured-graph batch bit-equal to the plain
        per-instance form over 20 generations, and 3 members (smallest,
        median, largest n) equal to the oracle.
  * C5  the 200-request served mix on a Poisson trace (500 req/s) with the
        200-generation deadline: every result equal to its solo run, and two
        requests (a TSP and a knapsack one) equal to the oracle's best.

The C2 shape (n = 1000, P = 1024, 4 islands) is covered in
test_gpu_parity.py::test_full_size_tsp_islands_parity.
"""
import os
import sys
import time

import numpy as np
import pytest

import brkga_amd
import pyoracle
from brkga_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (instance set and request mix of the benchmark legs)

pytestmark = pytest.mark.gpu


def compare(pop, ost, ctx):
    r = pop.read()
    assert np.array_equal(r["keys"], ost.keys), f"keys differ ({ctx})"
    assert np.array_equal(r["fitness"], ost.fitness), f"fitness differs ({ctx})"
    assert np.array_equal(r["order"], ost.order), f"rank order differs ({ctx})"


@pytest.mark.parametrize("graph", [False, True])
def test_c1_tsp52_p256_500_generations(engine, graph):
    print(f"C1 TSP n=52 P=256 500 generations graph={graph}", flush=True)
    xy = workloads.tsp_coords(52, 52, 1000.0)
    gp = brkga_amd.Problem.tsp_coords(engine, xy)
    op = pyoracle.Problem.tsp_coords(xy)
    P, E, M, seed = 256, 51, 25, 0xC1C1
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E, M, 0.7, 1, 0, 0, seed))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, 0.7, seed=seed))
    compare(pop, ost, "after init")
    for block in range(10):
        pop.evolve(50, graph=graph)
        for _ in range(50):
            ost.step()
        compare(pop, ost, f"after {50 * (block + 1)} generations")
    assert pop.generation == 500
    fit, keys, tour = pop.best()
    assert sorted(tour.tolist()) == list(range(52))
    assert fit == ost.fitness.min()


def test_c3a_greedy_knapsack_n500_p2048(engine):
    print("C3a greedy knapsack n=500 P=2048 (one population)", flush=True)
    w, v, cap = workloads.knapsack(500, 500)
    gp = brkga_amd.Problem.knapsack(engine, w, v, cap, greedy=True)
    op = pyoracle.Problem.knapsack(w, v, cap, True)
    P, E, M, seed = 2048, 409, 204, 0xC3A
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E, M, 0.7, 1, 0, 0, seed))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, 0.7, seed=seed, maximize=1))
    compare(pop, ost, "after init")
    for t in range(30):
        pop.evolve(1, graph=(t % 3 == 2))
        ost.step()
        if t % 5 == 4 or t < 3:
            compare(pop, ost, f"after step {t}")
    compare(pop, ost, "after 30 generations")
    # integer fitness: the best is the exact value of its decoded item set
    fit, keys, sel = pop.best()
    assert fit == float(int((sel.astype(np.int64) * v).sum()))
    assert int((sel.astype(np.int64) * w).sum()) <= cap
    assert fit == ost.fitness.max()


def test_c3b_tspj_n200_p1024(engine):
    print("C3b TSPJ n=200 P=1024", flush=True)
    travel, jobs = workloads.tspj(200, 200)
    gp = brkga_amd.Problem.tspj(engine, travel, jobs)
    op = pyoracle.Problem.tspj(travel, jobs)
    P, E, M, seed = 1024, 204, 102, 0xC3B
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E, M, 0.75, 1, 0, 0, seed))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, 0.75, seed=seed))
    compare(pop, ost, "after init")
    for t in range(20):
        pop.evolve(1, graph=(t >= 10))
        ost.step()
        if t % 4 == 3:
            compare(pop, ost, f"after step {t}")
    compare(pop, ost, "after 20 generations")


def test_c4_batched_512_instances(engine):
    """The benchmark's batched instance set: graph replay of the batched
    launches equals the plain per-instance form (every member's keys,
    fitness and rank order), and three members equal the oracle."""
    G = 20
    sizes = [int(n) for n in workloads.log_uniform_sizes(bench.BATCH_COUNT, bench.BATCH_N_LO, bench.BATCH_N_HI)]
    print(f"C4 batched {len(sizes)} instances, sum n = {sum(sizes)}, {G} generations", flush=True)
    xys = [workloads.tsp_coords(1000 + i, n, 1000.0) for i, n in enumerate(sizes)]
    probs = [brkga_amd.Problem.tsp_coords(engine, xy) for xy in xys]
    params = [brkga_amd.Params(bench.BATCH_POP, bench.BATCH_ELITE, bench.BATCH_MUTANTS, bench.BATCH_RHO, 1, 0, 0,
                               bench.SEED + i) for i in range(len(sizes))]
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize("launch")
    batch.autotune(reps=1)
    batch.initialize("launch")
    batch.evolve(G, "graph")
    graph_states = [batch.member(i).read() for i in range(len(sizes))]
    assert all(batch.member(i).generation == G for i in range(len(sizes)))
    print("C4 graph form done; plain form", flush=True)
    batch.initialize("plain")
    batch.evolve(G, "plain")
    for i in range(len(sizes)):
        got = batch.member(i).read()
        for key in ("keys", "fitness", "order"):
            assert np.array_equal(got[key], graph_states[i][key]), f"member {i} (n={sizes[i]}): {key} differs"
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for i in picks:
        print(f"C4 member {i} (n={sizes[i]}) against the oracle", flush=True)
        ost = pyoracle.BrkgaState(pyoracle.Problem.tsp_coords(xys[i]),
                                  pyoracle.params(bench.BATCH_POP, bench.BATCH_ELITE, bench.BATCH_MUTANTS,
                                                  bench.BATCH_RHO, seed=params[i].seed))
        for _ in range(G):
            ost.step()
        want = graph_states[i]
        assert np.array_equal(want["keys"], ost.keys), f"member {i}: keys differ from the oracle"
        assert np.array_equal(want["fitness"], ost.fitness), f"member {i}: fitness differs from the oracle"
        assert np.array_equal(want["order"], ost.order), f"member {i}: rank order differs from the oracle"
    batch.close()
    for p in probs:
        p.close()


def test_c5_served_poisson_mix(engine):
    """The benchmark's 200-request mix, arriving as a Poisson stream of 500
    requests per second: every request's best fitness and decoded solution
    equal its solo run (plain launches), and the first TSP and the first
    knapsack request of the mix equal the oracle's best after 200
    generations."""
    count, rate = 200, 500.0
    reqs = bench.serve_requests(brkga_amd, workloads, count)
    arrivals = workloads.poisson_arrivals(count, rate)
    print(f"C5 served mix: {count} requests at {rate} req/s", flush=True)
    srv = brkga_amd.Server(engine, hbm_budget_bytes=8 << 30, ticks_per_round=4)
    ids = []
    t0 = time.perf_counter()
    for (kind, _, data, prm), ta in zip(reqs, arrivals):
        wait = ta - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(bench.serve_problem(brkga_amd, engine, kind, data), prm, bench.SERVE_GENS,
                              own_problem=True))
    served = [srv.wait(r, timeout_ms=120000) for r in ids]
    st = srv.stats()
    srv.stop()
    assert st["completed"] == count and st["peak_bytes"] <= st["budget_bytes"]
    print("C5 solo runs", flush=True)
    for i, ((kind, n, data, prm), (fit, dec, info)) in enumerate(zip(reqs, served)):
        assert info["status"] == 0 and info["generations"] == bench.SERVE_GENS, f"request {i}"
        gp = bench.serve_problem(brkga_amd, engine, kind, data)
        pop = brkga_amd.Population(gp, prm)
        pop.initialize()
        pop.evolve(bench.SERVE_GENS)
        sfit, _, sdec = pop.best()
        pop.close()
        gp.close()
        assert np.array_equal(np.asarray(fit), np.asarray(sfit)), f"request {i} ({kind} n={n}): best differs"
        assert np.array_equal(dec, sdec), f"request {i} ({kind} n={n}): decoded solution differs"
    for want_kind in ("tsp", "knapsack"):
        i = next(j for j, r in enumerate(reqs) if r[0] == want_kind and r[1] <= 600)
        kind, n, data, prm = reqs[i]
        print(f"C5 request {i} ({kind} n={n}) against the oracle", flush=True)
        if kind == "tsp":
            op = pyoracle.Problem.tsp_coords(data)
        else:
            w, v, cap = data
            op = pyoracle.Problem.knapsack(w, v, cap, True)
        ost = pyoracle.BrkgaState(op, pyoracle.params(prm.population_size, prm.elite_size, prm.mutant_size,
                                                      prm.elite_prob, seed=prm.seed, maximize=op.maximize))
        for _ in range(bench.SERVE_GENS):
            ost.step()
        best = ost.fitness.max() if op.maximize else ost.fitness.min()
        assert served[i][0] == best, f"request {i} ({kind} n={n}): best differs from the oracle"
