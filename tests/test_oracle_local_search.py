"""TSPJ local search of the oracle (oracle/oracle.c orc_tspj_local_search).

Pins: the move semantics are the definition of 2-opt on the decoded tour
(reverse the tour between two drawn positions), so the accepted candidate's
decoded tour must equal the original tour with the drawn segments reversed,
the drawn positions coming from the engine's Philox stream (numpy
restatement in brkga_amd.workloads); acceptance is strict improvement;
job keys and the multiset of city keys never change; the generation
schedule and the top-k choice follow solver.hpp:458-462,566.
"""
import numpy as np

import pyoracle
from brkga_amd import workloads

TAG_LSEARCH = 5


def words(seed, island, g, rank, count):
    """Philox words 0..count-1 of counter (island, g, rank, LSEARCH << 24 | block)."""
    out = []
    for blk in range((count + 3) // 4):
        r = workloads.philox4x32_10(np.uint32(island), np.uint32(g), np.uint32(rank),
                                    np.uint32((TAG_LSEARCH << 24) | blk), seed & 0xFFFFFFFF, seed >> 32)
        out.extend(int(x) for x in r)
    return out[:count]


def setup(n=20, P=32, E=8, M=4, islands=1, seed=77, ls=(2, 4, 6)):
    travel, jobs = workloads.tspj(3, n)
    prob = pyoracle.Problem.tspj(travel, jobs)
    q = pyoracle.params(P, E, M, 0.75, islands, 0, 0, seed, 0, *ls)
    st = pyoracle.BrkgaState(prob, q)
    return prob, q, st


def test_moves_zero_and_off_schedule_change_nothing():
    prob, q, st = setup()
    for g, qq in ((2, pyoracle.params(32, 8, 4, 0.75, 1, 0, 0, 77, 0, 2, 4, 0)),   # no moves
                  (3, q),                                                        # 3 % 2 != 0
                  (0, q)):                                                       # generation 0
        k, f, p = st.keys.copy(), st.fitness.copy(), st.order.copy()
        pyoracle.tspj_local_search(prob, qq, g, k, f, p)
        assert np.array_equal(k, st.keys) and np.array_equal(f, st.fitness) and np.array_equal(p, st.order)


def test_accepted_candidates_are_drawn_two_opt_reversals():
    n = 20
    prob, q, st = setup(n=n, ls=(2, 4, 6))
    seen_accept = 0
    for trial in range(12):
        st.step()
        g = 2 * (trial + 1)
        k, f, p = st.keys.copy(), st.fitness.copy(), st.order.copy()
        pyoracle.tspj_local_search(prob, q, g, k, f, p)
        for r in range(4):
            row = int(st.order[r])
            before, after = st.keys[row], k[row]
            assert np.array_equal(after[n:], before[n:]), "job keys changed"
            assert np.array_equal(np.sort(after[:n]), np.sort(before[:n])), "city key multiset changed"
            assert f[row] <= st.fitness[row]
            if np.array_equal(after, before):
                continue
            seen_accept += 1
            assert f[row] < st.fitness[row], "kept without strict improvement"
            tour = list(np.argsort(before[:n], kind="stable"))
            w = words(77, 0, g, r, 12)
            for m in range(6):
                x, y = w[2 * m] % n, w[2 * m + 1] % n
                x, y = min(x, y), max(x, y)
                if y - x >= 2:
                    tour[x:y + 1] = tour[x:y + 1][::-1]
            assert list(np.argsort(after[:n], kind="stable")) == tour
            fit, _ = prob.evaluate(after.reshape(1, -1))
            assert fit[0] == f[row]
        # order is the stable partition of the updated fitness
        assert np.array_equal(p, pyoracle.partition(f, 1, 32, 0))
        st.keys, st.fitness, st.order = k, f, p
    assert seen_accept > 0, "no move was ever accepted: the check above never ran"


def test_step_applies_search_on_schedule_only():
    prob, q, st = setup(P=32, E=8, islands=2, ls=(3, 2, 5))
    prob2, q2, st2 = setup(P=32, E=8, islands=2, ls=(0, 0, 0))
    for t in range(7):
        st.step()
        st2.step()
        g = t + 1
        if g < 3:  # identical until the first search at generation 3
            assert np.array_equal(st.keys, st2.keys)
    # the search only ever lowers the island minimum at the generation it runs
    assert st.fitness.min() <= st2.fitness.max()
