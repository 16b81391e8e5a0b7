"""Pins the oracle's NSGA-II diversity injection (ensure_minimum_diversity,
core/population.hpp:130-182, run after the generations core/solver.hpp:812-814
names).

The first pass is pinned by the three properties that determine a greedy
first-fit selection in index order (kept indices increase; kept members are
pairwise at least 0.001 apart; every dropped member is closer than 0.001 to
some kept member before it), checked with FP64 numpy distances on random
tiny populations full of near duplicates, plus constructed cases: exact
duplicate groups, the 0.001 boundary on both sides in one and two
dimensions, and a chain whose middle point is close to both ends (greedy
keeps both ends; comparing with every earlier member, kept or not, would
drop the far end).  The injection is pinned at the 70% boundary (7 unique
of 10 keeps the population, 6 of 10 replaces it) and by its output: kept
rows in order, new rows from the Philox DIVERSITY stream, objectives from
the evaluator, ranks and crowding from the (separately pinned) front and
crowding functions.
"""
import math

import numpy as np
import pytest

import pyoracle
from brkga_amd import workloads

EPS = 0.001


def dist(a, b):
    return math.sqrt(sum((float(x) - float(y)) ** 2 for x, y in zip(a, b)))


def check_greedy(obj, kept):
    kept = list(kept)
    assert kept == sorted(kept) and len(set(kept)) == len(kept)
    for i, a in enumerate(kept):
        for b in kept[i + 1:]:
            assert dist(obj[a], obj[b]) >= EPS
    ks = set(kept)
    for i in range(obj.shape[0]):
        if i not in ks:
            assert any(k < i and dist(obj[i], obj[k]) < EPS for k in kept), i


def test_unique_properties_random_tiny():
    g = np.random.default_rng(7)
    for trial in range(300):
        n = int(g.integers(1, 14))
        base = g.random((int(g.integers(1, 5)), 2)) * 0.004
        obj = (base[g.integers(0, base.shape[0], n)] + g.normal(0, 0.0006, (n, 2))).astype(np.float32)
        check_greedy(obj, pyoracle.diversity_unique(obj))


def test_unique_exact_groups_and_chain():
    obj = np.array([[1, 2], [3, 4], [1, 2], [3, 4], [5, 5], [1, 2]], np.float32)
    assert pyoracle.diversity_unique(obj).tolist() == [0, 1, 4]
    chain = np.array([[0.0, 0.0], [0.0006, 0.0], [0.0012, 0.0]], np.float32)
    assert pyoracle.diversity_unique(chain).tolist() == [0, 2]


def test_unique_boundary():
    below = np.nextafter(np.float32(0.001), np.float32(0))
    while float(below) >= EPS:
        below = np.nextafter(below, np.float32(0))
    above = np.float32(0.001)
    assert float(below) < EPS <= float(above)
    for x, kept in ((below, [0]), (above, [0, 1])):
        obj = np.array([[0, 0], [x, 0]], np.float32)
        assert pyoracle.diversity_unique(obj).tolist() == kept
    # diagonal: distance sqrt(2) a on each side of 0.001
    a = np.float32(EPS / math.sqrt(2))
    lo = a
    while dist((lo, lo), (0, 0)) >= EPS:
        lo = np.nextafter(lo, np.float32(0))
    hi = np.nextafter(lo, np.float32(1))
    assert dist((lo, lo), (0, 0)) < EPS <= dist((hi, hi), (0, 0))
    assert pyoracle.diversity_unique(np.array([[0, 0], [lo, lo]], np.float32)).tolist() == [0]
    assert pyoracle.diversity_unique(np.array([[0, 0], [hi, hi]], np.float32)).tolist() == [0, 1]


def population_with_unique(prob, N, U, seed):
    """N members of which the first U distinct objective points are unique
    and the rest duplicate them."""
    keys = pyoracle.init_keys(seed, 0, N, prob.chrom_len)
    obj, _ = prob.evaluate(keys)
    src = np.arange(N) % U
    keys, obj = keys[src].copy(), obj[src].copy()
    assert len(pyoracle.diversity_unique(obj)) == U
    rank = pyoracle.nds_ranks(obj)
    return keys, obj, rank, pyoracle.crowding(obj, rank)


@pytest.mark.parametrize("U,injects", [(7, False), (6, True), (10, False), (1, True)])
def test_injection_at_the_70_percent_boundary(U, injects):
    prob = pyoracle.Problem.bitsp_coords(*workloads.bitsp_coords(4, 30))
    N, t, seed = 10, 25, 99
    q = pyoracle.params(N, 0, 0, 0.7, seed=seed, div_interval=25)
    keys, obj, rank, crowd = population_with_unique(prob, N, U, seed=U)
    k2, o2, r2, c2 = keys.copy(), obj.copy(), rank.copy(), crowd.copy()
    ran = pyoracle.lib().orc_nsga2_diversity(pyoracle.ctypes.byref(prob.c), pyoracle.ctypes.byref(q), t,
                                            pyoracle._f(k2), pyoracle._f(o2), pyoracle._i(r2), pyoracle._f(c2))
    assert bool(ran) == injects
    if not injects:
        for a, b in ((k2, keys), (o2, obj), (r2, rank), (c2, crowd)):
            assert np.array_equal(a, b)
        return
    kept = pyoracle.diversity_unique(obj)
    assert np.array_equal(k2[:U], keys[kept]) and np.array_equal(o2[:U], obj[kept])
    L = prob.chrom_len
    for r in range(U, N):
        words = [pyoracle.philox_word(seed, 0, t + 1, r, 6, j) for j in range(L)]
        assert np.array_equal(k2[r], np.array([pyoracle.u32_to_unit(w) for w in words], np.float32))
    want_obj, _ = prob.evaluate(k2)
    assert np.array_equal(o2, want_obj)
    assert np.array_equal(r2, pyoracle.nds_ranks(o2))
    assert np.array_equal(c2, pyoracle.crowding(o2, r2))


def test_step_schedule():
    """Injections happen only after steps t with t % interval == 0, and a
    run with the option off never injects."""
    prob = pyoracle.Problem.bitsp_coords(*workloads.bitsp_coords(5, 4))  # 4 cities: few distinct tours
    for interval in (0, 3, 25):
        st = pyoracle.Nsga2State(prob, pyoracle.params(32, 0, 0, 0.7, seed=5, div_interval=interval))
        hits = []
        for t in range(12):
            st.step()
            if st.injected:
                hits.append(t)
        assert all(interval and t % interval == 0 for t in hits)
        if interval:
            assert 0 in hits  # 32 members over 3 distinct tours
        else:
            assert not hits
