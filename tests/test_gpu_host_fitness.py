"""Host-evaluated problems (a config whose fitness_function runs on the host
and has no engine decoder, core/config.hpp:33-39 evaluated as
core/population.hpp:626-653 does): keys built and partitioned on the device,
fitness computed by the caller between brkga_population_host_begin and
_end.

The bar: a host fitness function that computes exactly what a device
decoder computes gives a trajectory bit-identical to the device-decoded
run, generation for generation (keys, fitness, rank orders; NSGA-II ranks
and crowding), with islands, migration and maximisation."""
import numpy as np
import pytest

import brkga_amd
from brkga_amd import workloads

pytestmark = pytest.mark.gpu


def greedy_knapsack(w, v, cap):
    """Greedy fill in ascending key order (stable by item index): integer
    value of the items that fit, as a host fitness function would do it."""
    w = np.asarray(w, np.int64)
    v = np.asarray(v, np.int64)

    def fitness(keys):
        out = np.empty(keys.shape[0], np.float32)
        for r, row in enumerate(keys):
            left, val = cap, 0
            for i in np.argsort(row, kind="stable"):
                if w[i] <= left:
                    left -= w[i]
                    val += v[i]
            out[r] = val
        return out
    return fitness


def same(a, b, keys, ctx):
    for k in keys:
        assert np.array_equal(a[k], b[k]), f"{k} differs ({ctx})"


@pytest.mark.parametrize("P,E,M,I,iv,k", [(64, 12, 6, 1, 0, 0), (4 * 48, 4 * 8, 4 * 6, 4, 2, 3)])
def test_host_knapsack_matches_device_greedy(engine, P, E, M, I, iv, k):
    n = 120
    w, v, cap = workloads.knapsack(7, n)
    native = brkga_amd.Problem.knapsack(engine, w, v, cap, greedy=True)
    host = brkga_amd.Problem.host(engine, n, 1, maximize=True)
    assert host.kind == brkga_amd.HOST and host.maximize == 1
    prm = brkga_amd.Params(P, E, M, 0.7, I, iv, k, 0xF17 + I)
    a = brkga_amd.Population(native, prm)
    b = brkga_amd.Population(host, prm)
    fit = greedy_knapsack(w, v, cap)
    a.initialize()
    b.host_generation(fit, initialize=True)
    same(a.read(), b.read(), ("keys", "fitness", "order"), "generation 0")
    for t in range(8):
        a.evolve(1)
        keys, f, needs = b.host_begin()
        idx = np.nonzero(needs)[0]
        # copied rows (elite, migrants) come back with their fitness
        want = a.read()
        assert np.array_equal(keys, want["keys"]), f"keys differ at step {t}"
        assert np.array_equal(f[~needs, 0], want["fitness"][~needs]), f"copied fitness differs at step {t}"
        assert needs.sum() < P
        f[idx, 0] = fit(keys[idx])
        b.host_end(f)
        same(want, b.read(), ("keys", "fitness", "order"), f"step {t}")
    assert b.generation == 8
    bf, bk, _ = b.best()
    af, ak, _ = a.best()
    assert bf == af and np.array_equal(bk, ak)


def test_host_nsga2_matches_device_bitsp(engine):
    """Bi-objective host problem whose objectives are computed through the
    device evaluator of a bi-objective TSP (brkga_evaluate_host): same NSGA-II
    trajectory as the device-decoded population."""
    a_xy, b_xy = workloads.bitsp_coords(3, 25)
    native = brkga_amd.Problem.bitsp_coords(engine, a_xy, b_xy)
    host = brkga_amd.Problem.host(engine, native.chrom_len, 2)
    prm = brkga_amd.Params(48, 0, 0, 0.7, 1, 0, 0, 0x2B)
    a = brkga_amd.Population(native, prm)
    b = brkga_amd.Population(host, prm)
    objectives = lambda keys: native.evaluate(keys)[0]
    a.initialize()
    b.host_generation(objectives, initialize=True)
    for t in range(6):
        a.evolve(1)
        b.host_generation(objectives)
        same(a.read(), b.read(), ("keys", "fitness", "rank", "crowding"), f"step {t}")


def test_host_problem_errors(engine):
    host = brkga_amd.Problem.host(engine, 30, 1)
    pop = brkga_amd.Population(host, brkga_amd.Params(32, 6, 4, 0.7, 1, 0, 0, 1))
    for call in (lambda: pop.initialize(), lambda: pop.evolve(1), lambda: pop.evolve(1, graph=True),
                 lambda: pop.host_end(np.zeros(32, np.float32)),
                 lambda: host.evaluate(np.zeros((2, 30), np.float32)),
                 lambda: brkga_amd.Batch(engine, [host], [brkga_amd.Params(32, 6, 4)]),
                 lambda: brkga_amd.Problem.host(engine, 30, 3),
                 lambda: brkga_amd.Problem.host(engine, 30, 2, maximize=True)):
        with pytest.raises(brkga_amd.BrkgaError) as e:
            call()
        assert e.value.status == brkga_amd.ERR_INVALID_ARG
    pop.host_begin(initialize=True)
    with pytest.raises(brkga_amd.BrkgaError):
        pop.host_begin()  # a generation is already open
    pop.host_end(np.ones(32, np.float32))
    mo = brkga_amd.Problem.host(engine, 30, 2)
    with pytest.raises(brkga_amd.BrkgaError) as e:
        brkga_amd.Population(mo, brkga_amd.Params(32, 0, 0, 0.7, 1, 0, 0, 1, diversity_interval=25))
    assert e.value.status == brkga_amd.ERR_INVALID_ARG
