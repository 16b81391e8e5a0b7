"""Step launch shapes against the oracle: the multi-child step (2 or 4
children of one island per workgroup, the island ranked once per
workgroup) and the one-child step with two keys per thread: keys, fitness and
rank orders bit-exact, including migration generations (the preceding
island ranked by a second group), islands of one wave, copies decoded
alongside built children, and rows that take the bounded long-bucket path
in some groups only."""
import numpy as np
import pytest

import brkga_amd
import pyoracle
from cases import make_problem
from conftest import random_keys

pytestmark = pytest.mark.gpu

CASES = [
    # n, islands, Pi, elite/island, mutants/island, interval, k, generations, block, children
    (1000, 4, 256, 38, 38, 3, 2, 8, 256, 4),
    (1000, 4, 256, 38, 38, 3, 2, 8, 256, 2),
    (500, 2, 128, 20, 12, 2, 3, 7, 128, 4),
    (240, 3, 64, 10, 6, 2, 4, 9, 64, 4),
    (60, 2, 64, 9, 7, 2, 2, 6, 64, 2),
    (1800, 2, 512, 60, 40, 2, 2, 5, 512, 2),
    # one child per workgroup, two keys per thread (float2 columns)
    (1000, 4, 256, 38, 38, 3, 2, 8, 512, 1),
    (501, 2, 256, 30, 20, 2, 2, 6, 512, 1),
    (2000, 2, 128, 20, 12, 2, 2, 5, 1024, 1),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}-I{c[1]}-Pi{c[2]}-b{c[8]}-c{c[9]}")
def test_multi_child_matches_oracle(engine, case):
    n, I, Pi, E, M, iv, k, G, block, ch = case
    print(f"multi-child {case}", flush=True)
    gp, op = make_problem(engine, "tsp", n, inst=71)
    P, seed = I * Pi, 0x3C + n
    pop = brkga_amd.Population(gp, brkga_amd.Params(P, E * I, M * I, 0.7, I, iv, k, seed))
    pop.set_block(block)
    pop.set_children(ch)
    assert pop.step_shape() == (block, ch)
    pop.initialize()
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E * I, M * I, 0.7, I, iv, k, seed))
    for t in range(G):
        pop.evolve(1, graph=(t % 2 == 1))
        ost.step()
        r = pop.read()
        assert np.array_equal(r["keys"], ost.keys), f"keys differ after step {t}"
        assert np.array_equal(r["fitness"], ost.fitness), f"fitness differs after step {t}"
        assert np.array_equal(r["order"], ost.order), f"rank order differs after step {t}"


def test_multi_child_long_buckets(engine):
    """Written rows of equal keys next to random rows: some groups of a
    workgroup take the sorted path, all groups take it together."""
    n, P, I = 1000, 512, 2
    gp, op = make_problem(engine, "tsp", n, inst=72)
    prm = brkga_amd.Params(P, 80, 60, 0.7, I, 2, 2, 0xE1)
    pop = brkga_amd.Population(gp, prm)
    pop.set_block(256)
    pop.set_children(4)
    keys = random_keys(9, P, n)
    keys[::5] = 0.5
    pop.write(keys)
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, 80, 60, 0.7, I, 2, 2, 0xE1))
    fit, _ = op.evaluate(keys)
    ost.keys, ost.fitness, ost.order, ost.t = keys.copy(), fit, pyoracle.partition(fit, I, P // I, 0), 0
    for t in range(3):
        pop.evolve(1)
        ost.step()
        r = pop.read()
        for key, want in (("keys", ost.keys), ("fitness", ost.fitness), ("order", ost.order)):
            assert np.array_equal(r[key], want), f"{key} differs after step {t}"


def test_children_fall_back_when_not_applicable(engine):
    """Kinds or shapes the multi-child step does not cover run the
    one-child step (same trajectory) and report 1 child per workgroup."""
    gp, _ = make_problem(engine, "knap_greedy", 300, inst=73)
    pop = brkga_amd.Population(gp, brkga_amd.Params(128, 20, 10, 0.7, 2, 0, 0, 5))
    pop.set_block(64)
    pop.set_children(4)
    assert pop.step_shape() == (64, 1)
    with pytest.raises(brkga_amd.BrkgaError):
        pop.set_children(3)
