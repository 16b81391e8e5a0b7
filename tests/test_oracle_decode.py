"""Pins the oracle's decoders and fitness functions.

Anchors: TSP configs/tsp_config.hpp:180-225,359-371; knapsack
configs/knapsack_config.hpp:43-75; TSPJ configs/tspj_config.hpp:555-642;
bi-objective TSP configs/multi_tsp_config.hpp:54-103; dominance
core/individual.hpp:99-114.  Pins: numpy's stable argsort, brute force on
tiny inputs, closed forms, FP64 recomputation (1e-6 relative) and the golden
vectors computed by the reference's own code (tests/golden/*_ref.npz).
"""
import itertools
import math
import os

import numpy as np
import pytest

import pyoracle
from conftest import random_keys

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


# ---------------------------------------------------------------- argsort --

@pytest.mark.parametrize("n", [1, 2, 3, 52, 127, 128, 129, 1000, 5000])
def test_argsort_matches_numpy_stable(n):
    for ties in (False, True):
        k = random_keys(n, 3, n, ties=ties)
        for row in k:
            assert np.array_equal(pyoracle.argsort_keys(row), np.argsort(row, kind="stable"))


def test_argsort_brute_force_tiny():
    # every key vector over {0, 0.25, 0.5} for n <= 5: the result must be the
    # unique permutation along which (key, index) strictly increases
    for n in range(1, 6):
        for vals in itertools.product([0.0, 0.25, 0.5], repeat=n):
            k = np.array(vals, np.float32)
            p = pyoracle.argsort_keys(k)
            assert sorted(p.tolist()) == list(range(n))
            seq = [(k[i], i) for i in p]
            assert all(seq[i] < seq[i + 1] for i in range(n - 1))


def test_argsort_signed_zero_and_extremes():
    k = np.array([0.5, -0.0, 0.0, 1.5, -2.0, 0.0], np.float32)
    # -0.0 == 0.0 compares equal, so index order decides between them
    assert pyoracle.argsort_keys(k).tolist() == [4, 1, 2, 5, 0, 3]


def test_argsort_nan_last_matches_numpy():
    """NaN keys sort after every number, in index order (numpy's stable
    argsort puts NaN last the same way); rows of equal keys, keys outside
    [0, 1) and NaN mixed in."""
    g = np.random.default_rng(11)
    rows = [np.array([0.5, np.nan, 0.25, np.nan, -1.0, 2.0, 0.25], np.float32),
            np.full(300, np.nan, np.float32),
            np.full(300, 0.375, np.float32),
            (g.integers(0, 100, 2000) + g.integers(0, 4, 2000) * 0.25).astype(np.float32)]
    mixed = g.random(2000).astype(np.float32)
    mixed[g.integers(0, 2000, 50)] = np.nan
    mixed[:7] = -0.0
    rows.append(mixed)
    for row in rows:
        assert np.array_equal(pyoracle.argsort_keys(row), np.argsort(row, kind="stable"))


# -------------------------------------------------------------------- TSP --

def tour_length_fp64(xy, perm):
    xy = xy.astype(np.float64)
    a = xy[perm]
    b = xy[np.roll(perm, -1)]
    return float(np.sqrt(((a - b) ** 2).sum(axis=1)).sum())


@pytest.mark.parametrize("n", [1, 2, 3, 52, 129, 1000])
def test_tsp_fitness_vs_fp64(n):
    g = np.random.default_rng(n)
    xy = (g.random((n, 2)) * 10000).astype(np.float32)
    prob = pyoracle.Problem.tsp_coords(xy)
    keys = random_keys(n + 1, 8, n)
    fit, perm = prob.evaluate(keys)
    for i in range(8):
        assert np.array_equal(perm[i], np.argsort(keys[i], kind="stable"))
        ref = tour_length_fp64(xy, perm[i])
        if ref == 0:
            assert fit[i] == 0
        else:
            assert abs(fit[i] - ref) / ref <= 1e-6


def test_tsp_regular_polygon_closed_form():
    for n, r in ((3, 1.0), (52, 500.0), (1000, 4000.0)):
        th = 2 * np.pi * np.arange(n) / n
        xy = np.stack([2 * r + r * np.cos(th), 2 * r + r * np.sin(th)], 1).astype(np.float32)
        keys = (np.arange(n, dtype=np.float32) / np.float32(n)).reshape(1, n)
        fit, perm = pyoracle.Problem.tsp_coords(xy).evaluate(keys)
        assert perm[0].tolist() == list(range(n))
        exact = 2 * n * r * math.sin(math.pi / n)
        # FP32 coordinates carry ~2e-7 relative rounding, far below the 1/n
        # relative error of a dropped or doubled edge
        assert abs(fit[0] - exact) / exact < 1e-4
        ref = tour_length_fp64(xy, perm[0])
        assert abs(fit[0] - ref) / ref <= 1e-6


def test_tsp_edge_formula_is_unfused_fp32():
    g = np.random.default_rng(5)
    xy = (g.random((200, 2)) * 10000).astype(np.float32)
    for a, b in g.integers(0, 200, size=(500, 2)):
        dx = np.float32(xy[a, 0] - xy[b, 0])
        dy = np.float32(xy[a, 1] - xy[b, 1])
        want = np.sqrt(np.float32(np.float32(dx * dx) + np.float32(dy * dy)))
        assert np.float32(pyoracle.lib().orc_tsp_edge_coords(xy.ctypes.data_as(pyoracle._PF), int(a), int(b))) == want


def test_tsp_matrix_kind_equals_coords_kind_on_euclidean_matrix():
    from brkga_amd.workloads import euclid_matrix

    g = np.random.default_rng(9)
    xy = (g.random((60, 2)) * 1000).astype(np.float32)
    keys = random_keys(1, 16, 60)
    f1, p1 = pyoracle.Problem.tsp_coords(xy).evaluate(keys)
    f2, p2 = pyoracle.Problem.tsp_matrix(euclid_matrix(xy)).evaluate(keys)
    assert np.array_equal(p1, p2)
    # the shifts differ (diameter bound vs matrix maximum), so compare to FP64
    for i in range(16):
        ref = tour_length_fp64(xy, p1[i])
        assert abs(f1[i] - ref) / ref <= 1e-6 and abs(f2[i] - ref) / ref <= 1e-6


def test_fixed_shift_rule():
    assert pyoracle.fixed_shift(0.0) == 30
    assert pyoracle.fixed_shift(1.0) == 47       # 1 < 2^1
    assert pyoracle.fixed_shift(14142.2) == 34   # < 2^14
    assert pyoracle.fixed_shift(2.0 ** 20) == 27
    assert pyoracle.fixed_shift(1e-30) == 60


# --------------------------------------------------------------- knapsack --

def _load_cases(name):
    z = np.load(os.path.join(GOLD, name))
    cnt = int(z["count"])
    keys = [k.rsplit("_", 1)[0] for k in z.files if k != "count"]
    return [{k: z[f"{k}_{i}"] for k in set(keys)} for i in range(cnt)]


def test_knapsack_threshold_matches_reference_golden():
    for c in _load_cases("knapsack_threshold_ref.npz"):
        prob = pyoracle.Problem.knapsack(c["w"], c["v"], int(c["cap"]), greedy=False)
        fit, sel = prob.evaluate(c["keys"])
        assert np.array_equal(fit, c["fitness"])
        assert np.array_equal(sel, c["selection"])


def test_knapsack_threshold_matches_compiled_reference():
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    g = np.random.default_rng(77)
    for n in (3, 40, 300):
        w = g.integers(1, 101, n).astype(np.int32)
        v = (w + g.integers(0, 11, n)).astype(np.int32)
        cap = int(w.sum()) // 3
        keys = random_keys(n, 50, n)
        fit, sel = pyoracle.Problem.knapsack(w, v, cap, greedy=False).evaluate(keys)
        rf = np.empty(50, np.float32)
        rs = np.empty((50, n), np.int32)
        wf, vf = w.astype(np.float32), v.astype(np.float32)
        ref.ref_knapsack_threshold(wf.ctypes.data_as(pyoracle._PF), vf.ctypes.data_as(pyoracle._PF), n,
                                   float(cap), keys.ctypes.data_as(pyoracle._PF), 50,
                                   rf.ctypes.data_as(pyoracle._PF), rs.ctypes.data_as(pyoracle._PI))
        assert np.array_equal(fit, rf) and np.array_equal(sel, rs)


def test_knapsack_greedy_properties():
    g = np.random.default_rng(12)
    for n in (1, 5, 12, 500, 2000):
        w = g.integers(1, 101, n).astype(np.int32)
        v = (w + g.integers(0, 11, n)).astype(np.int32)
        cap = int(w.sum()) // 2
        keys = random_keys(n + 7, 6, n, ties=(n == 12))
        fit, sel = pyoracle.Problem.knapsack(w, v, cap).evaluate(keys)
        for r in range(6):
            order = np.argsort(keys[r], kind="stable")
            taken = sel[r].astype(bool)
            assert int(w[taken].sum()) <= cap
            assert fit[r] == float(v[taken].sum())
            # walking the key order, an item is skipped only if it did not fit
            rem = cap
            for i in order:
                if taken[i]:
                    rem -= int(w[i])
                else:
                    assert int(w[i]) > rem
            # every skipped item exceeds the final slack
            assert all(int(w[i]) > rem for i in np.nonzero(~taken)[0])


def test_knapsack_greedy_closed_forms():
    w = np.array([5, 3, 9, 1], np.int32)
    v = np.array([10, 4, 11, 2], np.int32)
    keys = np.array([[0.4, 0.1, 0.2, 0.3]], np.float32)  # order 1, 2, 3, 0
    assert pyoracle.Problem.knapsack(w, v, 100).evaluate(keys)[0][0] == 27.0   # everything fits
    assert pyoracle.Problem.knapsack(w, v, 0).evaluate(keys)[0][0] == 0.0      # nothing fits
    # cap 12: take 1 (3), take 2 (9) -> rem 0, skip 3, skip 0
    f, s = pyoracle.Problem.knapsack(w, v, 12).evaluate(keys)
    assert f[0] == 15.0 and s[0].tolist() == [0, 1, 1, 0]
    # cap 9: take 1 (3), skip 2 (9 > 6), take 3 (1), skip 0 (5 > 5? no: 5 <= 5) -> take 0
    f, s = pyoracle.Problem.knapsack(w, v, 9).evaluate(keys)
    assert s[0].tolist() == [1, 1, 0, 1] and f[0] == 16.0


# ------------------------------------------------------------------- TSPJ --

def tspj_fp64(travel, jobs, tour, job):
    t = travel.astype(np.float64)
    j = jobs.astype(np.float64)
    cur, time, best = 0, 0.0, 0.0
    for c, k in zip(tour, job):
        time += t[cur, c]
        best = max(best, time + j[c, k])
        cur = c
    return max(best, time + t[cur, 0])


def test_tspj_closed_form_one_and_two_cities():
    travel = np.array([[0, 3], [4, 0]], np.float32)
    jobs = np.array([[0, 0], [0, 10]], np.float32)
    f, d = pyoracle.Problem.tspj(travel, jobs).evaluate(np.array([[0.3, 0.6]], np.float32))
    assert d[0].tolist() == [1, 1] and f[0] == 13.0            # max(3+4, 3+10)
    jobs[1, 1] = 1.0
    assert pyoracle.Problem.tspj(travel, jobs).evaluate(np.array([[0.3, 0.6]], np.float32))[0][0] == 7.0
    travel = np.array([[0, 1, 2], [1, 0, 5], [2, 5, 0]], np.float32)
    jobs = np.array([[0, 0, 0], [0, 7, 1], [0, 2, 9]], np.float32)
    # city keys (0.9, 0.1) -> tour 2, 1; job keys (0.2, 0.8) -> jobs 1, 2
    f, d = pyoracle.Problem.tspj(travel, jobs).evaluate(np.array([[0.9, 0.1, 0.2, 0.8]], np.float32))
    assert d[0].tolist() == [2, 1, 1, 2]
    # arrive 2 at t=2, job 1 -> 4; arrive 1 at t=7, job 2 -> 8; back at 8
    assert f[0] == 8.0


def test_tspj_vs_fp64():
    from brkga_amd.workloads import tspj

    travel, jobs = tspj(3, 200)
    prob = pyoracle.Problem.tspj(travel, jobs)
    keys = random_keys(4, 16, 400)
    f, d = prob.evaluate(keys)
    for r in range(16):
        assert np.array_equal(d[r, :200], np.argsort(keys[r, :200], kind="stable") + 1)
        assert np.array_equal(d[r, 200:], np.argsort(keys[r, 200:], kind="stable") + 1)
        ref = tspj_fp64(travel, jobs, d[r, :200], d[r, 200:])
        assert abs(f[r] - ref) / ref <= 1e-6


# ---------------------------------------------------------- bi-objective --

def test_multitsp_matches_reference_golden():
    for c in _load_cases("multitsp_ref.npz"):
        prob = pyoracle.Problem.bitsp_matrix(c["m0"], c["m1"])
        obj, perm = prob.evaluate(c["keys"])
        assert np.array_equal(perm, c["perm"])
        # the reference sums sequentially in FP32; both must agree with FP64
        for r in range(obj.shape[0]):
            for m, mat in enumerate((c["m0"], c["m1"])):
                tour = [0] + (perm[r] + 1).tolist() + [0]
                ref = sum(float(mat[a, b]) for a, b in zip(tour[:-1], tour[1:]))
                assert abs(obj[r, m] - ref) <= 1e-6 * ref
                assert abs(c["objectives"][r, m] - ref) <= 2e-6 * ref


def test_bitsp_coords_equals_matrix_form():
    from brkga_amd.workloads import bitsp_coords, euclid_matrix

    a, b = bitsp_coords(1, 80)
    keys = random_keys(2, 8, 80)
    o1, p1 = pyoracle.Problem.bitsp_coords(a, b).evaluate(keys)
    o2, p2 = pyoracle.Problem.bitsp_matrix(euclid_matrix(a), euclid_matrix(b)).evaluate(keys)
    assert np.array_equal(p1, p2)
    assert np.allclose(o1, o2, rtol=1e-6, atol=0)


def test_dominance_matches_reference_golden():
    z = np.load(os.path.join(GOLD, "dominance_ref.npz"))
    obj, dom = z["objectives"], z["dominates"]
    n = obj.shape[0]
    # ranks from the oracle: i dominated by j implies rank[j] < rank[i]
    rank = pyoracle.nds_ranks(obj)
    for i in range(n):
        for j in range(n):
            if dom[j, i]:
                assert rank[j] < rank[i]
    # rank 0 is exactly the set nobody dominates
    assert set(np.nonzero(rank == 0)[0]) == set(np.nonzero(dom.sum(axis=0) == 0)[0])
