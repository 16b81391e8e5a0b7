"""Pins the oracle's partition and BRKGA generation step.

Anchors: partition core/population.hpp:660-682 (comparator order; stable, as
thrust::sort_by_key at core/solver.hpp:416-419); generation step
core/cuda_kernels.cuh:88-138 (elite copy, elite parent from [0,E),
non-elite parent from [E,P), coin < rho, mutants); islands and ring
migration core/solver.hpp:378-476,517-558.  Pins: numpy stable sorts and
slot-by-slot invariants recomputed from the Philox stream.
"""
import numpy as np
import pytest

import pyoracle
from brkga_amd import workloads
from conftest import random_keys


def test_partition_matches_numpy_stable():
    g = np.random.default_rng(1)
    for seg in (1, 7, 256, 1000):
        f = g.integers(0, 20, size=4 * seg).astype(np.float32)  # heavy ties
        for mx in (0, 1):
            perm = pyoracle.partition(f, 4, seg, mx).reshape(4, seg)
            for s in range(4):
                x = f[s * seg:(s + 1) * seg]
                want = np.argsort(-x if mx else x, kind="stable")
                assert np.array_equal(perm[s], want)


def test_partition_special_values():
    f = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, 1.0], np.float32)
    assert pyoracle.partition(f, 1, 6, 0).tolist() == [4, 1, 2, 5, 3, 0]
    assert pyoracle.partition(f, 1, 6, 1).tolist() == [3, 5, 1, 2, 4, 0]


def _tsp(n=40, seed=0):
    return pyoracle.Problem.tsp_coords(workloads.tsp_coords(seed, n))


def _unit(seed, stream, gen, slot, tag, j):
    return np.float32(pyoracle.u32_to_unit(pyoracle.philox_word(seed, stream, gen, slot, tag, j)))


def test_init_population():
    prob = _tsp()
    q = pyoracle.params(64, 12, 6, 0.7, islands=2, seed=99)
    st = pyoracle.BrkgaState(prob, q)
    for i in range(2):
        assert np.array_equal(st.keys[i * 32:(i + 1) * 32], pyoracle.init_keys(99, i, 32, 40))
    fit, _ = prob.evaluate(st.keys)
    assert np.array_equal(st.fitness, fit)
    assert np.array_equal(st.order, pyoracle.partition(fit, 2, 32, 0))


@pytest.mark.parametrize("islands,interval,k", [(1, 0, 0), (4, 3, 2)])
def test_generation_step_invariants(islands, interval, k):
    prob = _tsp(30, 1)
    P, E, M, rho, seed = 4 * 24, 4 * 5, 4 * 4, 0.7, 1234
    q = pyoracle.params(P, E, M, rho, islands=islands, mig_interval=interval, mig_count=k, seed=seed)
    st = pyoracle.BrkgaState(prob, q)
    Pi, Ei, Mi = P // islands, E // islands, M // islands
    for t in range(7):
        keys0, order0 = st.keys.copy(), st.order.copy()
        st.step()
        g = t + 1
        migrate = islands > 1 and interval > 0 and t > 0 and t % interval == 0
        for i in range(islands):
            base = i * Pi
            for s in range(Pi):
                row = st.keys[base + s]
                if migrate and s >= Pi - k:
                    src = (i - 1) % islands
                    assert np.array_equal(row, keys0[src * Pi + order0[src * Pi + s - (Pi - k)]])
                elif s < Ei:
                    assert np.array_equal(row, keys0[base + order0[base + s]])
                elif s >= Pi - Mi:
                    want = [_unit(seed, i, g, s, pyoracle.TAG_MUTANT, j) for j in range(30)]
                    assert np.array_equal(row, np.array(want, np.float32))
                else:
                    u0 = pyoracle.philox_word(seed, i, g, s, pyoracle.TAG_PARENTS, 0)
                    u1 = pyoracle.philox_word(seed, i, g, s, pyoracle.TAG_PARENTS, 1)
                    pe = keys0[base + order0[base + u0 % Ei]]
                    pn = keys0[base + order0[base + Ei + u1 % (Pi - Ei)]]
                    for j in range(30):
                        c = _unit(seed, i, g, s, pyoracle.TAG_COINS, j)
                        assert row[j] == (pe[j] if c < np.float32(rho) else pn[j])
        fit, _ = prob.evaluate(st.keys)
        assert np.array_equal(st.fitness, fit)
        assert np.array_equal(st.order, pyoracle.partition(fit, islands, Pi, 0))


def test_elite_parent_frequency_is_rho():
    prob = _tsp(200, 2)
    q = pyoracle.params(64, 8, 8, 0.7, seed=5)
    st = pyoracle.BrkgaState(prob, q)
    keys0, order0 = st.keys.copy(), st.order.copy()
    st.step()
    hits = total = 0
    for s in range(8, 56):
        u0 = pyoracle.philox_word(5, 0, 1, s, pyoracle.TAG_PARENTS, 0)
        u1 = pyoracle.philox_word(5, 0, 1, s, pyoracle.TAG_PARENTS, 1)
        pe = keys0[order0[u0 % 8]]
        pn = keys0[order0[8 + u1 % 56]]
        diff = pe != pn
        hits += int((st.keys[s][diff] == pe[diff]).sum())
        total += int(diff.sum())
    assert abs(hits / total - 0.7) < 0.02


def test_determinism_and_seed_sensitivity():
    prob = _tsp(50, 3)
    q = pyoracle.params(32, 6, 4, 0.7, seed=7)
    a, b = pyoracle.BrkgaState(prob, q), pyoracle.BrkgaState(prob, q)
    for _ in range(5):
        a.step()
        b.step()
    assert np.array_equal(a.keys, b.keys) and np.array_equal(a.fitness, b.fitness)
    c = pyoracle.BrkgaState(prob, pyoracle.params(32, 6, 4, 0.7, seed=8))
    assert not np.array_equal(a.keys, c.keys)


def test_brkga_improves_tsp():
    prob = _tsp(52, 4)
    q = pyoracle.params(256, 51, 25, 0.7, seed=11)
    st = pyoracle.BrkgaState(prob, q)
    f0 = st.fitness.min()
    for _ in range(150):
        st.step()
    # elites survive, so the best never gets worse; 150 generations cut it a lot
    assert st.fitness.min() < 0.6 * f0


def test_knapsack_maximises():
    w, v, cap = workloads.knapsack(0, 100)
    prob = pyoracle.Problem.knapsack(w, v, cap)
    q = pyoracle.params(128, 25, 12, 0.7, seed=3, maximize=1)
    st = pyoracle.BrkgaState(prob, q)
    best = [st.fitness[st.order[0]]]
    for _ in range(30):
        st.step()
        assert st.fitness[st.order[0]] == st.fitness.max()
        best.append(st.fitness[st.order[0]])
    assert all(b1 >= b0 for b0, b1 in zip(best, best[1:]))
