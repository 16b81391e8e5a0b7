"""GPU parity of the two island-ranking paths: the step ranking its parent
island inside every workgroup (pow2(Pi) <= workgroup size) and the
partition launch after each step (islands wider than the workgroup).

Both must give the oracle's trajectory bit for bit: keys, fitness and rank
orders.  The cases cover islands that are not powers of two (padded
networks), islands of one wave and of several waves (LDS exchange steps),
the widest in-step islands (1024), islands wider than the workgroup's
staged rank order (2560, 3000: the step reads the partition's order from
global memory), ring migration on the generations the cases run through, maximisation with many equal fitness values (knapsack), and a mix
of single launches and graph replays.
"""
import numpy as np
import pytest

import brkga_amd
import pyoracle
from cases import make_problem

pytestmark = pytest.mark.gpu

RANK_CASES = [
    # kind, n, Pi, islands, elite/island, mutants/island, interval, k, generations, blocks
    ("tsp", 50, 37, 1, 7, 5, 0, 0, 5, (64, 128)),
    ("tsp", 80, 100, 3, 17, 11, 2, 3, 6, (64, 128, 256)),
    ("knap_greedy", 200, 200, 2, 30, 20, 3, 2, 7, (64, 128, 256)),
    ("knap_threshold", 60, 64, 2, 10, 6, 2, 2, 5, (64, 128)),
    ("tsp", 40, 512, 2, 60, 40, 2, 4, 4, (256, 512)),
    ("tsp_matrix", 30, 1000, 1, 100, 80, 0, 0, 3, (512, 1024)),
    ("tsp", 30, 2560, 1, 300, 200, 0, 0, 3, (64, 256)),
    ("tsp", 25, 3000, 2, 400, 300, 2, 5, 4, (128,)),
]


@pytest.mark.parametrize("case", RANK_CASES, ids=lambda c: f"{c[0]}-n{c[1]}-Pi{c[2]}-I{c[3]}")
def test_in_step_and_partition_ranking_match_oracle(engine, case):
    kind, n, Pi, I, E, M, iv, k, G, blocks = case
    print(f"island rank {case}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=7)
    P, seed = Pi * I, 0xA11CE + Pi
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E * I, M * I, 0.7, I, iv, k, seed, op.maximize))
    for t in range(G):
        ost.step()
    want = (ost.keys.copy(), ost.fitness.copy(), ost.order.copy())
    for blk in blocks:
        pop = brkga_amd.Population(gp, brkga_amd.Params(P, E * I, M * I, 0.7, I, iv, k, seed))
        pop.set_block(blk)
        pop.initialize()
        pop.evolve(1)
        pop.evolve(G - 2, graph=True)
        pop.evolve(1)
        r = pop.read()
        ctx = f"block {blk} ({'in-step' if Pi <= blk else 'partition'} ranking)"
        assert pop.generation == G, ctx
        assert np.array_equal(r["keys"], want[0]), f"keys differ, {ctx}"
        assert np.array_equal(r["fitness"], want[1]), f"fitness differs, {ctx}"
        assert np.array_equal(r["order"], want[2]), f"rank order differs, {ctx}"
        pop.close()
