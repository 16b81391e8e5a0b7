"""Checkpoint / resume (brkga_population_restore): a population checkpointed
after some generations and restored into a fresh population continues
bit-identically to the uninterrupted run, across islands, migration, the
TSPJ local search, NSGA-II with the diversity injection, graph replays and
host-evaluated problems."""
import numpy as np
import pytest

import brkga_amd
from cases import make_problem

pytestmark = pytest.mark.gpu

CASES = [
    # kind, n, Params, stop, total
    ("tsp", 200, brkga_amd.Params(4 * 64, 4 * 10, 4 * 8, 0.7, 4, 5, 2, 0x5A), 13, 31),
    ("knap_greedy", 300, brkga_amd.Params(96, 16, 10, 0.7, 1, 0, 0, 0x5B), 7, 20),
    ("tspj", 40, brkga_amd.Params(64, 12, 6, 0.75, 2, 3, 2, 0x5C, 2, 4, 5), 9, 18),
    ("bitsp", 5, brkga_amd.Params(48, 0, 0, 0.7, 1, 0, 0, 0x5D, diversity_interval=3), 10, 22),
    ("zdt1", 12, brkga_amd.Params(64, 0, 0, 0.6, 1, 0, 0, 0x5E), 4, 12),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-n{c[1]}-at{c[3]}")
@pytest.mark.parametrize("graph", [False, True])
def test_restore_continues_bit_identically(engine, case, graph):
    kind, n, prm, stop, total = case
    print(f"restore {kind} n={n} at {stop} of {total} graph={graph}", flush=True)
    gp, _ = make_problem(engine, kind, n, inst=61)
    a = brkga_amd.Population(gp, prm)
    a.initialize()
    a.evolve(total)
    want = a.read()
    b = brkga_amd.Population(gp, prm)
    b.initialize()
    b.evolve(stop)
    ck = b.checkpoint()
    assert ck["generation"] == stop
    b.close()
    c = brkga_amd.Population(gp, prm)
    c.restore(ck)
    assert c.generation == stop
    for key in ck:
        if key != "generation":
            assert np.array_equal(c.read()[key], ck[key]), f"{key} not restored"
    c.evolve(total - stop, graph=graph)
    assert c.generation == total
    got = c.read()
    for key in want:
        assert np.array_equal(got[key], want[key]), f"{key} differs after resuming"


def test_restore_host_evaluated(engine):
    w = np.arange(1, 41, dtype=np.int32)
    v = w + 3
    cap = int(w.sum()) // 3

    def fit(keys):
        out = []
        for row in keys:
            left, val = cap, 0
            for i in np.argsort(row, kind="stable"):
                if w[i] <= left:
                    left -= w[i]
                    val += v[i]
            out.append(val)
        return np.array(out, np.float32)

    hp = brkga_amd.Problem.host(engine, 40, 1, maximize=True)
    prm = brkga_amd.Params(64, 10, 6, 0.7, 2, 2, 2, 0x5F)
    a = brkga_amd.Population(hp, prm)
    a.host_generation(fit, initialize=True)
    for _ in range(10):
        a.host_generation(fit)
    want = a.read()
    b = brkga_amd.Population(hp, prm)
    b.host_generation(fit, initialize=True)
    for _ in range(4):
        b.host_generation(fit)
    ck = b.checkpoint()
    c = brkga_amd.Population(hp, prm)
    c.restore(ck)
    for _ in range(6):
        c.host_generation(fit)
    got = c.read()
    for key in want:
        assert np.array_equal(got[key], want[key]), key


def test_restore_rejects_bad_arguments(engine):
    gp, _ = make_problem(engine, "bitsp", 10, inst=62)
    pop = brkga_amd.Population(gp, brkga_amd.Params(32, 0, 0, 0.7, 1, 0, 0, 1))
    pop.initialize()
    ck = pop.checkpoint()
    L = brkga_amd.lib()
    k = np.ascontiguousarray(ck["keys"])
    f = np.ascontiguousarray(ck["fitness"])
    fp = k.ctypes.data_as(brkga_amd._PF)
    assert L.brkga_population_restore(pop.handle, fp, f.ctypes.data_as(brkga_amd._PF), None, None, 0, None) == \
        brkga_amd.ERR_INVALID_ARG
    ck["generation"] = -1
    with pytest.raises(brkga_amd.BrkgaError):
        pop.restore(ck)
