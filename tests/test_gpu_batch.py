"""Batched form: each member of a batch evolves exactly as it does alone.

Members of mixed kinds and uneven sizes share the queue-driven step
launches, one partition launch and the NSGA-II launches; their keys,
fitness, rank orders (and NSGA-II ranks / crowding) must equal those of the
same population evolved alone through the single-population launches,
which test_gpu_parity.py pins to the oracle.  One member is also stepped
against the oracle directly.
"""
import numpy as np
import pytest
import torch

import brkga_amd
import pyoracle
from cases import make_problem

pytestmark = pytest.mark.gpu

MEMBERS = [
    # kind, n, P, E, M, rho, islands, migration interval, migrants
    ("tsp", 20, 64, 12, 6, 0.7, 1, 0, 0),
    ("tsp", 64, 64, 10, 10, 0.7, 2, 3, 2),
    ("tsp", 65, 128, 20, 12, 0.7, 4, 2, 1),
    ("tsp", 300, 64, 10, 10, 0.7, 1, 0, 0),
    ("tsp", 1500, 64, 10, 10, 0.7, 2, 4, 2),
    ("tsp", 2100, 32, 5, 5, 0.7, 1, 0, 0),
    ("tsp_matrix", 40, 48, 9, 5, 0.65, 1, 0, 0),
    ("knap_greedy", 300, 64, 12, 6, 0.7, 1, 0, 0),
    ("knap_greedy", 100, 64, 12, 6, 0.7, 1, 0, 0),
    ("knap_threshold", 120, 64, 12, 6, 0.7, 2, 2, 3),
    ("tspj", 60, 64, 12, 6, 0.75, 1, 0, 0),
    ("tspj_ls", 45, 64, 12, 6, 0.75, 2, 3, 2),
    ("bitsp", 30, 32, 0, 0, 0.7, 1, 0, 0),
    ("bitsp", 80, 64, 0, 0, 0.7, 1, 0, 0),
    ("cont_rastrigin", 30, 64, 12, 6, 0.7, 2, 2, 2),
    ("cont_rosenbrock", 200, 64, 12, 6, 0.7, 1, 0, 0),
    ("jobshop", 80, 64, 12, 6, 0.7, 1, 0, 0),
    ("vrp_days", 50, 64, 12, 6, 0.7, 1, 0, 0),
    ("zdt1", 30, 32, 0, 0, 0.5, 1, 0, 0),
    ("movrp", 31, 32, 0, 0, 0.5, 1, 0, 0),
]
NSGA_KINDS = ("bitsp", "zdt1", "movrp")


def build(engine):
    """tspj_ls: TSPJ with the local search every 2 generations (top 3 per
    island, 5 moves)."""
    probs, params = [], []
    for i, (kind, n, P, E, M, rho, I, iv, k) in enumerate(MEMBERS):
        ls = (2, 3, 5) if kind == "tspj_ls" else (0, 0, 0)
        gp, _ = make_problem(engine, "tspj" if kind == "tspj_ls" else kind, n, inst=100 + i)
        probs.append(gp)
        params.append(brkga_amd.Params(P, E, M, rho, I, iv, k, 0xBA7C4 + 7 * i, *ls))
    return probs, params


def solo_states(probs, params, G):
    out = []
    for gp, prm in zip(probs, params):
        pop = brkga_amd.Population(gp, prm)
        pop.initialize()
        pop.evolve(G)
        out.append(pop.read())
        pop.close()
    return out


def assert_same(batch, solo, G, ctx):
    for i, ref in enumerate(solo):
        m = batch.member(i)
        assert m.generation == G, f"member {i} at generation {m.generation} ({ctx})"
        got = m.read()
        for key in ref:
            assert np.array_equal(got[key], ref[key]), f"member {i} {MEMBERS[i][:2]}: {key} differs ({ctx})"


@pytest.mark.parametrize("mode", ["launch", "graph", "plain"])
def test_batch_members_match_solo(engine, mode):
    print(f"batch vs solo, mode={mode}", flush=True)
    probs, params = build(engine)
    G = 7
    solo = solo_states(probs, params, G)
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize(mode)
    if mode == "graph":
        # two graph lengths, the second replayed with eager work in between
        batch.evolve(3, mode)
        torch.cuda.synchronize()
        _ = batch.member(0).read()
        batch.evolve(2, mode)
        torch.cuda.synchronize()
        _ = batch.member(3).read()
        batch.evolve(2, mode)
    else:
        batch.evolve(G, mode)
    torch.cuda.synchronize()
    assert_same(batch, solo, G, mode)
    batch.close()


def test_batch_autotune_keeps_results(engine):
    print("batch autotune", flush=True)
    probs, params = build(engine)
    G = 5
    solo = solo_states(probs, params, G)
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize("launch")
    before = batch.buckets()
    batch.autotune(reps=1)
    after = batch.buckets()
    assert [(b["kind"], b["n_min"], b["n_max"], b["members"]) for b in before] == \
        [(b["kind"], b["n_min"], b["n_max"], b["members"]) for b in after]
    for b in after:
        assert b["block"] in (64, 128, 256, 512, 1024) and b["block"] * 16 >= b["n_max"]
        assert b["chunk"] >= 1
    assert sum(b["members"] for b in after) == sum(1 for m in MEMBERS if m[0] not in NSGA_KINDS)
    batch.evolve(G, "launch")
    torch.cuda.synchronize()
    assert_same(batch, solo, G, "after autotune")
    batch.close()


def test_batch_member_matches_oracle(engine):
    """Member 2 (TSP n=65, 4 islands, migration every 2) against the oracle,
    step by step, inside a batch with every other member."""
    print("batch member vs oracle", flush=True)
    probs, params = build(engine)
    kind, n, P, E, M, rho, I, iv, k = MEMBERS[2]
    _, op = make_problem(engine, kind, n, inst=102)
    q = params[2]
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, rho, I, iv, k, q.seed, op.maximize))
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize("launch")
    m = batch.member(2)
    for t in range(6):
        r = m.read()
        assert np.array_equal(r["keys"], ost.keys), f"keys differ at generation {t}"
        assert np.array_equal(r["fitness"], ost.fitness), f"fitness differs at generation {t}"
        assert np.array_equal(r["order"], ost.order), f"rank order differs at generation {t}"
        batch.evolve(1, "launch")
        ost.step()
    batch.close()


@pytest.mark.parametrize("lanes", ["1", "3"])
def test_batch_fixed_lanes(engine, lanes, monkeypatch):
    """BRKGA_FLEET_LANES fixes the stream count (no probing); results do not
    depend on it, and the count is reported."""
    print(f"batch with BRKGA_FLEET_LANES={lanes}", flush=True)
    monkeypatch.setenv("BRKGA_FLEET_LANES", lanes)
    probs, params = build(engine)
    G = 4
    solo = solo_states(probs, params, G)
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize("launch")
    batch.evolve(G, "launch")
    torch.cuda.synchronize()
    assert batch.lanes() == int(lanes)
    assert_same(batch, solo, G, f"{lanes} lanes")
    batch.close()


def test_batch_probed_lanes(engine, monkeypatch):
    """Without the variable the probes choose 1 to 4 lanes."""
    print("batch with probed lanes", flush=True)
    monkeypatch.delenv("BRKGA_FLEET_LANES", raising=False)
    probs, params = build(engine)
    batch = brkga_amd.Batch(engine, probs, params)
    batch.initialize("launch")
    batch.evolve(2, "launch")
    torch.cuda.synchronize()
    assert 1 <= batch.lanes() <= 4
    batch.close()
