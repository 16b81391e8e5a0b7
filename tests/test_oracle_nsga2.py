"""Pins the oracle's NSGA-II operators.

Anchors: dominance core/individual.hpp:99-114; fronts
core/population.hpp:334-369,409-440; crowding core/population.hpp:453-489;
survivors core/population.hpp:491-517; tournament / crossover / mutation
core/solver.hpp:1042-1091.  Pins: brute-force definitions (the front index
is the longest dominance chain ending at a point), front invariants, the
crowding closed form on collinear points, and the Philox stream.
"""
import numpy as np
import pytest

import pyoracle
from brkga_amd import workloads


def dom(a, b):
    return bool(np.all(a <= b) and np.any(a < b))


def brute_ranks(obj):
    n = obj.shape[0]
    rank = [None] * n
    order = sorted(range(n), key=lambda i: (obj[i, 0], obj[i, 1]))
    for i in order:  # dominators precede i in lexicographic order
        rank[i] = max([rank[j] + 1 for j in order if rank[j] is not None and dom(obj[j], obj[i])], default=0)
    return np.array(rank)


@pytest.mark.parametrize("ties", [False, True])
def test_ranks_match_longest_chain(ties):
    g = np.random.default_rng(3 + ties)
    for n in (1, 2, 5, 60, 300):
        obj = (g.integers(0, 7, (n, 2)) if ties else g.random((n, 2))).astype(np.float32)
        r = pyoracle.nds_ranks(obj)
        assert np.array_equal(r, brute_ranks(obj))
        for k in range(r.max() + 1):
            f = np.nonzero(r == k)[0]
            assert not any(dom(obj[a], obj[b]) for a in f for b in f)
            if k > 0:
                prev = np.nonzero(r == k - 1)[0]
                assert all(any(dom(obj[p], obj[b]) for p in prev) for b in f)


def test_crowding_collinear_closed_form():
    # one front on f1 = 10 - f0 with unequal spacing; index order is shuffled
    f0 = np.array([0.0, 1.0, 3.0, 6.0, 10.0], np.float32)
    perm = np.array([3, 0, 4, 1, 2])
    obj = np.stack([f0, 10 - f0], 1)[perm].astype(np.float32)
    rank = np.zeros(5, np.int32)
    c = pyoracle.crowding(obj, rank)
    want = {0.0: np.inf, 10.0: np.inf, 1.0: 2 * 3 / 10, 3.0: 2 * 5 / 10, 6.0: 2 * 7 / 10}
    for i in range(5):
        assert c[i] == pytest.approx(want[float(obj[i, 0])], rel=1e-6) or (np.isinf(c[i]) and np.isinf(want[float(obj[i, 0])]))


def test_crowding_small_fronts_and_zero_range():
    obj = np.array([[1, 1], [2, 0], [0, 2], [5, 5], [5, 5], [5, 5]], np.float32)
    r = pyoracle.nds_ranks(obj)
    assert r.tolist() == [0, 0, 0, 1, 1, 1]
    c = pyoracle.crowding(obj, r)
    assert np.isinf(c[1]) and np.isinf(c[2]) and c[0] == pytest.approx(2.0)
    # identical points: boundaries infinite, the middle gets nothing (range 0)
    assert np.isinf(c[3]) and np.isinf(c[5]) and c[4] == 0.0


def test_selection_fills_fronts_then_crowding():
    g = np.random.default_rng(8)
    obj = g.random((200, 2)).astype(np.float32)
    r = pyoracle.nds_ranks(obj)
    c = pyoracle.crowding(obj, r)
    order = pyoracle.nsga2_select(obj, r, c, 100)
    assert len(set(order.tolist())) == 100
    sel_r = r[order]
    assert np.all(np.diff(sel_r) >= 0)
    last = sel_r.max()
    assert all(((r < last) <= np.isin(np.arange(200), order)))
    cut = np.nonzero(r == last)[0]
    kept = [i for i in order if r[i] == last]
    dropped = [i for i in cut if i not in kept]
    if dropped:
        assert min(c[kept]) >= max(c[dropped])


def test_nsga2_step_determinism_and_progress():
    a, b = workloads.bitsp_coords(0, 40)
    prob = pyoracle.Problem.bitsp_coords(a, b)
    q = pyoracle.params(64, 0, 0, 0.7, seed=21)
    s1, s2 = pyoracle.Nsga2State(prob, q), pyoracle.Nsga2State(prob, q)
    start = s1.obj.sum(axis=1).min()
    for _ in range(40):
        s1.step()
        s2.step()
    assert np.array_equal(s1.keys, s2.keys) and np.array_equal(s1.obj, s2.obj)
    assert s1.obj.sum(axis=1).min() < start
    # survivors carry ranks / crowding recomputed over parents + children
    obj_check, _ = prob.evaluate(s1.keys)
    assert np.array_equal(obj_check, s1.obj)
    assert np.all(np.diff(s1.rank) >= 0)


def test_nsga2_offspring_stream():
    a, b = workloads.bitsp_coords(1, 12)
    prob = pyoracle.Problem.bitsp_coords(a, b)
    seed, N, L = 5, 16, 12
    q = pyoracle.params(N, 0, 0, 0.6, seed=seed)
    st = pyoracle.Nsga2State(prob, q)
    keys0, rank0, crowd0 = st.keys.copy(), st.rank.copy(), st.crowd.copy()
    st.step()

    def tourn(x, y):
        if rank0[x] < rank0[y]:
            return x
        if rank0[y] < rank0[x]:
            return y
        return x if crowd0[x] > crowd0[y] else y

    children = []
    for c in range(N):
        w = [pyoracle.philox_word(seed, 0, 1, c, pyoracle.TAG_PARENTS, i) for i in range(4)]
        p1, p2 = tourn(w[0] % N, w[1] % N), tourn(w[2] % N, w[3] % N)
        child = []
        for j in range(L):
            coin = pyoracle.u32_to_unit(pyoracle.philox_word(seed, 0, 1, c, pyoracle.TAG_COINS, j))
            x = keys0[p1, j] if np.float32(coin) < np.float32(0.6) else keys0[p2, j]
            mc = pyoracle.u32_to_unit(pyoracle.philox_word(seed, 0, 1, c, pyoracle.TAG_MUTCOIN, j))
            if np.float32(mc) < np.float32(5.0 / L):
                x = pyoracle.u32_to_unit(pyoracle.philox_word(seed, 0, 1, c, pyoracle.TAG_MUTANT, j))
            child.append(x)
        children.append(np.array(child, np.float32))
    pool = np.concatenate([keys0, np.stack(children)])
    # every survivor is a parent or one of the recomputed children
    for row in st.keys:
        assert any(np.array_equal(row, p) for p in pool)
