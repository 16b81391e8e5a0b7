"""GPU parity of the NSGA-II diversity injection (ensure_minimum_diversity,
core/population.hpp:130-182, scheduled as core/solver.hpp:812-814 does)
against the oracle: keys, objectives, front ranks and crowding after every
step, on problems whose populations collapse onto few objective points (the
injection fires) and on ones that stay diverse (it does not), through plain
launches, graph replays and the batched form."""
import numpy as np
import pytest

import brkga_amd
import pyoracle
from cases import make_problem

pytestmark = pytest.mark.gpu


def compare(pop, ost, ctx):
    r = pop.read()
    assert np.array_equal(r["keys"], ost.keys), f"keys differ {ctx}"
    assert np.array_equal(r["fitness"], ost.obj), f"objectives differ {ctx}"
    assert np.array_equal(r["rank"], ost.rank), f"ranks differ {ctx}"
    assert np.array_equal(r["crowding"], ost.crowd), f"crowding differs {ctx}"


CASES = [
    # kind, n, N, interval, generations, graph, injections at least
    ("bitsp", 4, 32, 3, 10, False, 2),
    ("bitsp", 5, 600, 1, 3, False, 3),     # N = 600: the selection counts dominance itself
    ("bitsp", 6, 128, 2, 9, True, 1),
    ("movrp", 5, 32, 2, 8, False, 2),
    ("bitsp", 30, 64, 25, 27, False, 0),   # the reference's schedule on a diverse population
    ("zdt1", 10, 64, 2, 6, True, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-n{c[1]}-N{c[2]}-every{c[3]}")
def test_diversity_trajectory_parity(engine, case):
    kind, n, N, interval, G, graph, min_inj = case
    print(f"diversity trajectory {case}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=31)
    seed = 0xD1 + n + N
    pop = brkga_amd.Population(gp, brkga_amd.Params(N, 0, 0, 0.7, 1, 0, 0, seed, diversity_interval=interval))
    pop.initialize()
    ost = pyoracle.Nsga2State(op, pyoracle.params(N, 0, 0, 0.7, seed=seed, div_interval=interval))
    compare(pop, ost, "after init")
    for t in range(G):
        pop.evolve(1, graph=graph and t >= 2)
        ost.step()
        compare(pop, ost, f"after step {t} (injected: {ost.injected})")
    assert ost.injections >= min_inj
    if min_inj == 0 and interval == 25:
        assert ost.injections == 0


def test_diversity_off_matches_plain_nsga2(engine):
    """diversity_interval = 0 leaves the trajectory of the plain NSGA-II."""
    gp, op = make_problem(engine, "bitsp", 4, inst=32)
    pop = brkga_amd.Population(gp, brkga_amd.Params(32, 0, 0, 0.7, 1, 0, 0, 5))
    pop.initialize()
    ost = pyoracle.Nsga2State(op, pyoracle.params(32, 0, 0, 0.7, seed=5))
    for t in range(5):
        pop.evolve(1)
        ost.step()
    compare(pop, ost, "diversity off")
    assert ost.injections == 0


def test_diversity_in_batch_matches_solo(engine):
    """Batched NSGA-II members with the injection on (and one with it off)
    evolve exactly as they do alone."""
    specs = [("bitsp", 4, 32, 2), ("bitsp", 6, 64, 3), ("movrp", 5, 32, 0), ("movrp", 4, 32, 2), ("bitsp", 5, 48, 1)]
    probs, params = [], []
    for i, (kind, n, N, iv) in enumerate(specs):
        gp, _ = make_problem(engine, kind, n, inst=40 + i)
        probs.append(gp)
        params.append(brkga_amd.Params(N, 0, 0, 0.7, 1, 0, 0, 0xBD + i, diversity_interval=iv))
    G = 7
    solo = []
    for gp, prm in zip(probs, params):
        p = brkga_amd.Population(gp, prm)
        p.initialize()
        p.evolve(G)
        solo.append(p.read())
        p.close()
    for mode in ("launch", "graph"):
        batch = brkga_amd.Batch(engine, probs, params)
        batch.initialize("launch")
        batch.evolve(G, mode)
        for i, ref in enumerate(solo):
            got = batch.member(i).read()
            for key in ref:
                assert np.array_equal(got[key], ref[key]), f"member {i} {specs[i]}: {key} differs ({mode})"
        batch.close()


def test_negative_interval_rejected(engine):
    gp, _ = make_problem(engine, "bitsp", 10, inst=33)
    with pytest.raises(brkga_amd.BrkgaError) as e:
        brkga_amd.Population(gp, brkga_amd.Params(32, 0, 0, 0.7, 1, 0, 0, 1, diversity_interval=-1))
    assert e.value.status == brkga_amd.ERR_INVALID_ARG
