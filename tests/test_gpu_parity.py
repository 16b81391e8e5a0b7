"""GPU parity: the HIP engine, called through its C ABI, against the oracle.

Bar: keys, rank orders, decoded permutations / item sets and integer
fitness bit-exact; FP32 tour lengths / makespans / objectives bit-exact to the
oracle (both use the same FP32 edges and order-independent fixed-point sums),
which itself stays within 1e-6 relative of FP64 (test_oracle_decode.py).
"""
import numpy as np
import pytest
import torch

import brkga_amd
import pyoracle
from brkga_amd import workloads
from cases import make_problem
from conftest import random_keys

pytestmark = pytest.mark.gpu


def gpu_evaluate(prob, keys):
    k = torch.from_numpy(keys).cuda()
    f = torch.empty((keys.shape[0], prob.num_obj), dtype=torch.float32, device="cuda")
    d = torch.empty(keys.shape, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    prob.evaluate_device(k.data_ptr(), f.data_ptr(), keys.shape[0])
    prob.decode_device(k.data_ptr(), d.data_ptr(), keys.shape[0])
    torch.cuda.synchronize()
    f = f.cpu().numpy()
    return (f[:, 0] if prob.num_obj == 1 else f), d.cpu().numpy()


EVAL_CASES = [("tsp", n) for n in (1, 2, 3, 52, 127, 128, 129, 1000, 5000, 12000)] + [
    ("tsp_matrix", 1), ("tsp_matrix", 60), ("tsp_matrix", 1000),
    ("knap_greedy", 1), ("knap_greedy", 63), ("knap_greedy", 500), ("knap_greedy", 2000),
    ("knap_threshold", 1), ("knap_threshold", 500), ("knap_threshold", 2000),
    ("tspj", 1), ("tspj", 2), ("tspj", 200), ("tspj", 1000),
    ("bitsp", 1), ("bitsp", 50), ("bitsp", 500), ("bitsp_matrix", 50)]


@pytest.mark.parametrize("kind,n", EVAL_CASES)
def test_evaluate_and_decode_parity(engine, kind, n):
    print(f"evaluate parity {kind} n={n}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=n)
    L = op.chrom_len
    for ties in (False, True):
        keys = random_keys(n * 3 + ties, 48, L, ties=ties)
        gf, gd = gpu_evaluate(gp, keys)
        of, od = op.evaluate(keys)
        assert np.array_equal(gd, od), f"decode mismatch ({kind}, n={n}, ties={ties})"
        assert np.array_equal(gf, of), f"fitness mismatch ({kind}, n={n}, ties={ties})"


@pytest.mark.parametrize("kind,n", [("tsp", 12000), ("tsp_matrix", 1000), ("knap_greedy", 2000)])
def test_problem_host_arrays_free_after_create(engine, kind, n):
    """Problem data go up asynchronously; the caller may overwrite its host
    arrays as soon as the create call returns (96 KB, 4 MB and two 8 KB
    arrays here)."""
    print(f"host arrays reused after create {kind} n={n}", flush=True)
    if kind == "tsp":
        xy = workloads.tsp_coords(n, n, 10000.0)
        op = pyoracle.Problem.tsp_coords(xy.copy())
        gp = brkga_amd.Problem.tsp_coords(engine, xy)
        arrays = [xy]
    elif kind == "tsp_matrix":
        d = workloads.euclid_matrix(workloads.tsp_coords(n, n))
        op = pyoracle.Problem.tsp_matrix(d.copy())
        gp = brkga_amd.Problem.tsp_matrix(engine, d)
        arrays = [d]
    else:
        w, v, cap = workloads.knapsack(n, n)
        w, v = np.ascontiguousarray(w, np.int32), np.ascontiguousarray(v, np.int32)
        op = pyoracle.Problem.knapsack(w.copy(), v.copy(), cap, True)
        gp = brkga_amd.Problem.knapsack(engine, w, v, cap, True)
        arrays = [w, v]
    for a in arrays:
        a[...] = 7
    keys = random_keys(n, 32, op.chrom_len)
    gf, gd = gpu_evaluate(gp, keys)
    of, od = op.evaluate(keys)
    assert np.array_equal(gd, od) and np.array_equal(gf, of)
    gp.close()


def test_problem_size_limits(engine):
    """Above 16000 genes, or beyond the decoder's LDS (TSP n = 16000 needs
    more than 160 KiB), creation fails with a clear error."""
    from brkga_amd import workloads
    for n in (16001, 16000):
        with pytest.raises(brkga_amd.BrkgaError) as e:
            brkga_amd.Problem.tsp_coords(engine, workloads.tsp_coords(3, n, 100.0))
        assert e.value.status == brkga_amd.ERR_INVALID_ARG
        assert ("problem size" in str(e.value)) if n > 16000 else ("LDS" in str(e.value))


def test_evaluate_rejects_wrong_chrom_len(engine):
    gp, _ = make_problem(engine, "tsp", 10)
    k = torch.zeros((2, 11), device="cuda")
    f = torch.zeros(2, device="cuda")
    st = brkga_amd.lib().brkga_evaluate(gp.handle, k.data_ptr(), f.data_ptr(), 2, 11, None)
    assert st == brkga_amd.ERR_INVALID_ARG


@pytest.mark.parametrize("case", ["cap0", "zero_weights", "all_fit", "exact_fit", "one_heavy"])
def test_greedy_knapsack_edges(engine, case):
    """Greedy fill boundaries: nothing fits, zero weights, everything fits,
    capacity equal to the total weight, one item heavier than the capacity."""
    print(f"greedy knapsack edge {case}", flush=True)
    g = np.random.default_rng(len(case))
    n = 300
    w = g.integers(1, 101, n).astype(np.int32)
    v = (w + g.integers(0, 11, n)).astype(np.int32)
    cap = int(w.sum()) // 2
    if case == "cap0":
        cap = 0
    elif case == "zero_weights":
        w[::3] = 0
    elif case == "all_fit":
        cap = int(w.sum()) * 2
    elif case == "exact_fit":
        cap = int(w.sum())
    elif case == "one_heavy":
        w[7] = 10 ** 6
        cap = int(w.sum()) - 10 ** 6 + 5
    gp = brkga_amd.Problem.knapsack(engine, w, v, cap, True)
    op = pyoracle.Problem.knapsack(w, v, cap, True)
    for ties in (False, True):
        keys = random_keys(77 + ties, 40, n, ties=ties)
        gf, gd = gpu_evaluate(gp, keys)
        of, od = op.evaluate(keys)
        assert np.array_equal(gd, od) and np.array_equal(gf, of), f"greedy knapsack {case} ties={ties}"


def test_greedy_knapsack_rejects_negative(engine):
    w = np.array([3, -1, 4], np.int32)
    v = np.array([1, 1, 1], np.int32)
    for ww, cap in ((w, 5), (np.abs(w), -1)):
        with pytest.raises(brkga_amd.BrkgaError) as e:
            brkga_amd.Problem.knapsack(engine, ww, v, cap, True)
        assert e.value.status == brkga_amd.ERR_INVALID_ARG


def compare_brkga(pop, ost, ctx):
    r = pop.read()
    assert np.array_equal(r["keys"], ost.keys), f"keys differ {ctx}"
    assert np.array_equal(r["fitness"], ost.fitness), f"fitness differs {ctx}"
    assert np.array_equal(r["order"], ost.order), f"rank order differs {ctx}"


BRKGA_CASES = [
    # kind, n, P, E, M, rho, islands, interval, k, generations
    ("tsp", 52, 64, 12, 6, 0.7, 1, 0, 0, 12),
    ("tsp", 100, 4 * 32, 4 * 6, 4 * 5, 0.7, 4, 3, 2, 10),
    ("tsp_matrix", 40, 48, 9, 5, 0.65, 1, 0, 0, 8),
    ("knap_greedy", 300, 64, 12, 6, 0.7, 1, 0, 0, 8),
    ("knap_threshold", 120, 64, 12, 6, 0.7, 2, 2, 3, 8),
    ("tspj", 60, 64, 12, 6, 0.75, 1, 0, 0, 8),
]


@pytest.mark.parametrize("case", BRKGA_CASES, ids=lambda c: f"{c[0]}-n{c[1]}-P{c[2]}-I{c[6]}")
def test_brkga_trajectory_parity(engine, case):
    kind, n, P, E, M, rho, I, iv, k, G = case
    print(f"brkga trajectory {case}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=1)
    seed = 0xC0FFEE + n
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E, M, rho, I, iv, k, seed))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, rho, I, iv, k, seed, op.maximize))
    compare_brkga(pop, ost, "after init")
    for t in range(G):
        pop.evolve(1)
        ost.step()
        compare_brkga(pop, ost, f"after step {t}")
    assert pop.generation == G


@pytest.mark.parametrize("graph", [False, True])
def test_tspj_local_search_trajectory_parity(engine, graph):
    """TSPJ with the local search every 2 generations on the top 4 of each
    of 2 islands, 6 random 2-opt moves each: bit-exact to the oracle."""
    print(f"tspj local search trajectory graph={graph}", flush=True)
    gp, op = make_problem(engine, "tspj", 40, inst=11)
    seed = 0x15EA
    pop = brkga_amd.Population(gp, brkga_amd.Params(64, 12, 6, 0.75, 2, 3, 2, seed, 2, 4, 6))
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(64, 12, 6, 0.75, 2, 3, 2, seed, 0, 2, 4, 6))
    compare_brkga(pop, ost, "after init")
    for t in range(9):
        pop.evolve(1, graph=graph)
        ost.step()
        compare_brkga(pop, ost, f"after step {t}")


def test_local_search_rejected_for_other_kinds(engine):
    gp, _ = make_problem(engine, "tsp", 30)
    with pytest.raises(brkga_amd.BrkgaError) as e:
        brkga_amd.Population(gp, brkga_amd.Params(32, 6, 4, 0.7, 1, 0, 0, 1, 2, 4, 5))
    assert e.value.status == brkga_amd.ERR_INVALID_ARG


def test_full_size_tsp_islands_parity(engine):
    """n=1000, P=1024 as 4 islands x 256, E=M=38 per island, ring migration
    of the top 2 every 50 generations: bit-exact over 52 generations, which
    include the migration at step 50."""
    print("full-size parity n=1000 P=1024", flush=True)
    xy = workloads.tsp_coords(0, 1000, 10000.0)
    gp = brkga_amd.Problem.tsp_coords(engine, xy)
    op = pyoracle.Problem.tsp_coords(xy)
    prm = brkga_amd.Params(1024, 152, 152, 0.7, 4, 50, 2, 2024)
    pop = brkga_amd.Population(gp, prm)
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(1024, 152, 152, 0.7, 4, 50, 2, 2024))
    compare_brkga(pop, ost, "init")
    for t in range(52):
        pop.evolve(1)
        ost.step()
        if t in (0, 1, 49, 50, 51):
            compare_brkga(pop, ost, f"step {t}")
    compare_brkga(pop, ost, "end")


def test_graph_replay_matches_plain(engine):
    gp, _ = make_problem(engine, "tsp", 200, inst=3)
    prm = brkga_amd.Params(4 * 64, 4 * 10, 4 * 10, 0.7, 4, 4, 2, 77)
    a = brkga_amd.Population(gp, prm)
    b = brkga_amd.Population(gp, prm)
    a.initialize()
    b.initialize()
    a.evolve(15)
    ra = a.read()
    # replay a 5-generation graph three times with eager work in between
    for _ in range(3):
        b.evolve(5, graph=True)
        torch.cuda.synchronize()
        _ = b.read()
    rb = b.read()
    for key in ("keys", "fitness", "order"):
        assert np.array_equal(ra[key], rb[key])
    assert b.generation == 15


@pytest.mark.parametrize("kind,n,P,E,M", [("tsp", 300, 64, 10, 8), ("knap_greedy", 500, 64, 10, 8),
                                          ("bitsp", 60, 64, 0, 0)])
def test_autotune_keeps_trajectory(engine, kind, n, P, E, M):
    """Workgroup-size tuning times launches that rewrite the same rows: the
    trajectory afterwards equals the untuned one."""
    print(f"autotune trajectory {kind} n={n}", flush=True)
    gp, _ = make_problem(engine, kind, n, inst=9)
    prm = brkga_amd.Params(P, E, M, 0.7, 1, 0, 0, 1234)
    a = brkga_amd.Population(gp, prm)
    b = brkga_amd.Population(gp, prm)
    a.initialize()
    b.initialize()
    a.evolve(2)
    b.evolve(2)
    blk = a.autotune(3)
    assert blk in (64, 128, 256, 512, 1024) and blk * 16 >= n
    a.evolve(4)
    a.evolve(3, graph=True)
    b.evolve(7)
    ra, rb = a.read(), b.read()
    for key in ra:
        assert np.array_equal(ra[key], rb[key]), key


def test_best_and_write(engine):
    gp, op = make_problem(engine, "tsp", 80, inst=4)
    prm = brkga_amd.Params(64, 12, 6, 0.7, 2, 0, 0, 5)
    pop = brkga_amd.Population(gp, prm)
    pop.initialize()
    pop.evolve(5)
    f, k, d = pop.best()
    of, od = op.evaluate(k.reshape(1, -1))
    assert f == of[0] and np.array_equal(d, od[0])
    r = pop.read()
    assert f == r["fitness"].min()
    # write: arbitrary keys become generation 0 of the oracle's partition
    keys = random_keys(9, 64, 80)
    pop.write(keys)
    r = pop.read()
    fit, _ = op.evaluate(keys)
    assert np.array_equal(r["keys"], keys) and np.array_equal(r["fitness"], fit)
    assert np.array_equal(r["order"], pyoracle.partition(fit, 2, 32, 0))
    assert pop.generation == 0


def test_invalid_params_rejected(engine):
    gp, _ = make_problem(engine, "tsp", 20)
    for prm in (brkga_amd.Params(100, 0, 10), brkga_amd.Params(100, 60, 50), brkga_amd.Params(10, 2, 1, 0.7, 3),
                brkga_amd.Params(100, 20, 10, 1.5), brkga_amd.Params(64, 8, 8, 0.7, 2, 5, 5)):
        with pytest.raises(brkga_amd.BrkgaError) as e:
            brkga_amd.Population(gp, prm)
        assert e.value.status == brkga_amd.ERR_INVALID_ARG


def compare_nsga(pop, ost, ctx):
    r = pop.read()
    assert np.array_equal(r["keys"], ost.keys), f"keys differ {ctx}"
    assert np.array_equal(r["fitness"], ost.obj), f"objectives differ {ctx}"
    assert np.array_equal(r["rank"], ost.rank), f"ranks differ {ctx}"
    assert np.array_equal(r["crowding"], ost.crowd), f"crowding differs {ctx}"


# N <= 512: the selection reads dominated-by bitsets; N = 600 (1200 rows)
# counts the dominance pairs inside the selection workgroup.
@pytest.mark.parametrize("n,N,G", [(20, 32, 10), (120, 512, 4), (50, 64, 30), (40, 600, 3)])
def test_nsga2_trajectory_parity(engine, n, N, G):
    print(f"nsga2 trajectory n={n} N={N}", flush=True)
    gp, op = make_problem(engine, "bitsp", n, inst=2)
    seed = 31337 + n
    pop = brkga_amd.Population(gp, brkga_amd.Params(N, 0, 0, 0.7, 1, 0, 0, seed))
    pop.initialize()
    ost = pyoracle.Nsga2State(op, pyoracle.params(N, 0, 0, 0.7, seed=seed))
    compare_nsga(pop, ost, "init")
    for t in range(G):
        pop.evolve(1)
        ost.step()
        compare_nsga(pop, ost, f"step {t}")


def test_set_block_and_profile(engine):
    """Any admissible workgroup size gives the same trajectory, whether the
    step ranks its parent island itself (islands of 128 at 256 and 1024
    threads) or a partition launch ranks it (64 threads); profile() runs
    real generations and returns positive kernel times (no partition time
    when the step ranks)."""
    print("set_block / profile", flush=True)
    gp, _ = make_problem(engine, "tsp", 300, inst=5)
    prm = brkga_amd.Params(256, 40, 24, 0.7, 2, 3, 2, 99)
    states = []
    for blk in (64, 256, 1024):
        pop = brkga_amd.Population(gp, prm)
        pop.set_block(blk)
        pop.initialize()
        pop.evolve(4)
        pop.evolve(3, graph=True)
        states.append(pop.read())
        if blk in (64, 256):
            g0 = pop.generation
            step_us, close_us = pop.profile(5)
            assert step_us > 0
            assert close_us > 0 if blk == 64 else close_us == 0
            assert pop.generation == g0 + 5
            pop.evolve(1)
            assert pop.generation == g0 + 6
        pop.close()
    for key in states[0]:
        assert np.array_equal(states[0][key], states[1][key]) and np.array_equal(states[0][key], states[2][key])
    pop = brkga_amd.Population(gp, prm)
    for bad in (32, 100, 2048):
        with pytest.raises(brkga_amd.BrkgaError):
            pop.set_block(bad)
    big, _ = make_problem(engine, "tsp", 2000, inst=6)
    with pytest.raises(brkga_amd.BrkgaError):
        brkga_amd.Population(big, prm).set_block(64)
    pop.close()
