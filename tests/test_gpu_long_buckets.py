"""GPU parity of the argsort on rows that put many keys in one bucket.

The decode's bucket sort ranks a key by scanning its bucket; rows whose keys
crowd one bucket (all keys equal, keys outside [0, 1), many NaN) instead
take the bounded path (the items are sorted by a bitonic network, each key
finds its own item).  Both paths must give the oracle's stable order
(ascending key, ties by index, NaN last) bit for bit, and a crowded row must
cost about as much as a sort, not the square of its length.
"""
import time

import numpy as np
import pytest
import torch

import brkga_amd
import pyoracle
from cases import make_problem
from conftest import random_keys

pytestmark = pytest.mark.gpu


def crowded_keys(pattern, rows, L, seed):
    g = np.random.default_rng(seed)
    if pattern == "equal":
        return np.full((rows, L), 0.375, np.float32)
    if pattern == "range100":
        return (g.integers(0, 100, (rows, L)) + g.integers(0, 4, (rows, L)) * 0.25).astype(np.float32)
    if pattern == "half_equal":
        k = random_keys(seed, rows, L)
        k[:, ::2] = 0.5
        return k
    if pattern == "nan":
        k = random_keys(seed, rows, L)
        k[g.random((rows, L)) < 0.3] = np.nan
        k[:, :5] = -0.0
        k[:, 5:9] = -3.0
        return k
    raise ValueError(pattern)


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


PATTERNS = ["equal", "range100", "half_equal", "nan"]


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("kind,n", [("tsp", 2000), ("tsp", 12000), ("tsp_matrix", 1500), ("knap_greedy", 2000),
                                    ("tspj", 2000), ("bitsp", 500)])
def test_crowded_rows_match_oracle(engine, kind, n, pattern):
    print(f"crowded rows {kind} n={n} {pattern}", flush=True)
    gp, op = make_problem(engine, kind, n, inst=21)
    keys = crowded_keys(pattern, 24, op.chrom_len, seed=n + len(pattern))
    gf, gd = gpu_evaluate(gp, keys)
    of, od = op.evaluate(keys)
    assert np.array_equal(gd, od), f"decode differs ({kind} n={n} {pattern})"
    assert np.array_equal(gf, of, equal_nan=True), f"fitness differs ({kind} n={n} {pattern})"


def test_crowded_rows_cost_a_sort(engine):
    """256 rows of 12000 equal keys decode in well under a second (a bucket
    scan would compare 144M item pairs per row)."""
    gp, _ = make_problem(engine, "tsp", 12000, inst=22)
    keys = np.full((256, 12000), 0.25, np.float32)
    k = torch.from_numpy(keys).cuda()
    d = torch.empty(keys.shape, dtype=torch.int32, device="cuda")
    gp.decode_device(k.data_ptr(), d.data_ptr(), 256)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gp.decode_device(k.data_ptr(), d.data_ptr(), 256)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"256 rows x 12000 equal keys: {dt * 1e3:.2f} ms", flush=True)
    assert dt < 0.5
    assert np.array_equal(d.cpu().numpy()[7], np.arange(12000, dtype=np.int32))


@pytest.mark.parametrize("n,block", [(1000, 256), (3000, 256), (700, 64)])
def test_step_decodes_written_equal_keys(engine, n, block):
    """A population written with equal keys: its children (elite copies and
    crossovers of equal rows) are rows of equal keys, decoded inside the step
    kernel (the register-key path at n <= 4 x block, the LDS path above)."""
    print(f"step on written equal keys n={n} block={block}", flush=True)
    gp, op = make_problem(engine, "tsp", n, inst=23)
    P, E, M, I, seed = 256, 40, 30, 2, 0xE0
    prm = brkga_amd.Params(P, E, M, 0.7, I, 2, 2, seed)
    pop = brkga_amd.Population(gp, prm)
    pop.set_block(block)
    keys = np.full((P, n), 0.5, np.float32)
    keys[1::3] = random_keys(5, len(keys[1::3]), n)
    pop.write(keys)
    ost = pyoracle.BrkgaState(op, pyoracle.params(P, E, M, 0.7, I, 2, 2, seed))
    fit, _ = op.evaluate(keys)
    ost.keys, ost.fitness, ost.order, ost.t = keys.copy(), fit, pyoracle.partition(fit, I, P // I, 0), 0
    for t in range(4):
        pop.evolve(1)
        ost.step()
        r = pop.read()
        assert np.array_equal(r["keys"], ost.keys), f"keys differ after step {t}"
        assert np.array_equal(r["fitness"], ost.fitness), f"fitness differs after step {t}"
        assert np.array_equal(r["order"], ost.order), f"rank order differs after step {t}"
