"""Writes the golden fixtures under tests/golden/ (run in the build container).

The reference ships no test vectors, so the fixtures come from:
  * philox_kat.json: Philox4x32-10 known-answer vectors published with
    Random123 (kat_vectors, philox4x32 10 rounds), entered by hand;
  * *_ref.npz: inputs and outputs of the reference's own CUDA-free code
    (KnapsackConfig, MultiTSPConfig, Individual::dominates), compiled from
    the reference sources into oracle/_ref/libref.so by `make -C oracle ref`.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402

PHILOX_KAT = [
    {"ctr": ["00000000", "00000000", "00000000", "00000000"], "key": ["00000000", "00000000"],
     "out": ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]},
    {"ctr": ["ffffffff", "ffffffff", "ffffffff", "ffffffff"], "key": ["ffffffff", "ffffffff"],
     "out": ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]},
    {"ctr": ["243f6a88", "85a308d3", "13198a2e", "03707344"], "key": ["a4093822", "299f31d0"],
     "out": ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]},
]


def _f(a):
    return a.ctypes.data_as(pyoracle._PF)


def _i(a):
    return a.ctypes.data_as(pyoracle._PI)


def knapsack_threshold(ref, out):
    g = np.random.default_rng(20240101)
    cases = []
    for n, rows in ((1, 8), (7, 32), (50, 64), (500, 16)):
        w = g.integers(1, 101, n).astype(np.float32)
        v = (w + g.integers(0, 11, n)).astype(np.float32)
        cap = np.float32(np.floor(0.5 * w.sum()))
        keys = (g.integers(0, 1 << 24, (rows, n)) * 2.0 ** -24).astype(np.float32)
        keys[0, :] = 0.5  # exactly at the threshold: not selected
        fit = np.empty(rows, np.float32)
        sel = np.empty((rows, n), np.int32)
        ref.ref_knapsack_threshold(_f(w), _f(v), n, float(cap), _f(keys), rows, _f(fit), _i(sel))
        cases.append(dict(w=w.astype(np.int32), v=v.astype(np.int32), cap=np.int64(cap), keys=keys, fitness=fit,
                          selection=sel))
    np.savez(out, **{f"{k}_{i}": val for i, c in enumerate(cases) for k, val in c.items()}, count=len(cases))


def multitsp(ref, out):
    g = np.random.default_rng(20240102)
    cases = []
    for n, rows in ((1, 4), (5, 16), (50, 32), (120, 8)):
        dim = n + 1
        xy0 = (g.random((dim, 2)) * 1000).astype(np.float32)
        xy1 = (g.random((dim, 2)) * 1000).astype(np.float32)
        m0 = _euclid(xy0)
        m1 = _euclid(xy1)
        keys = (g.integers(0, 1 << 24, (rows, n)) * 2.0 ** -24).astype(np.float32)
        if n >= 5:
            keys[0, : n // 2] = keys[0, 0]  # forced ties
        obj = np.empty((rows, 2), np.float32)
        perm = np.empty((rows, n), np.int32)
        ref.ref_multitsp(_f(m0), _f(m1), dim, _f(keys), rows, _f(obj), _i(perm))
        cases.append(dict(m0=m0, m1=m1, keys=keys, objectives=obj, perm=perm))
    np.savez(out, **{f"{k}_{i}": val for i, c in enumerate(cases) for k, val in c.items()}, count=len(cases))


def dominance(ref, out):
    g = np.random.default_rng(20240103)
    obj = g.integers(0, 6, (40, 2)).astype(np.float32)  # many ties
    res = np.empty((40, 40), np.int32)
    ref.ref_dominates(_f(obj), 40, 2, _i(res))
    np.savez(out, objectives=obj, dominates=res)


def loaders(ref, out):
    """The reference's own knapsack and bi-objective TSP file loaders on the
    fixture files in data/ (the TSP and TSPJ loaders need CUDA headers)."""
    data = os.path.join(HERE, "data")
    res = {}
    for name in ("knap_small.txt", "knap_cap.csv", "knap_nocap.csv"):
        w = np.zeros(64, np.float32)
        v = np.zeros(64, np.float32)
        cap = np.zeros(1, np.float32)
        cnt = ref.ref_knapsack_load(os.path.join(data, name).encode(), _f(w), _f(v), 64, _f(cap))
        key = name.replace(".", "_")
        res[key + "_w"] = w[:cnt]
        res[key + "_v"] = v[:cnt]
        res[key + "_cap"] = cap
    g = np.random.default_rng(20240104)
    n = 7  # tiny8.tsp: depot + 7 cities
    keys = (g.integers(0, 1 << 24, (24, n)) * 2.0 ** -24).astype(np.float32)
    keys[0] = np.arange(n, dtype=np.float32) / n  # identity tour
    obj = np.zeros((24, 2), np.float32)
    got = ref.ref_multitsp_load_objectives(os.path.join(data, "tiny8.tsp").encode(), _f(keys), 24, _f(obj))
    assert got == n
    res["tiny8_keys"] = keys
    res["tiny8_objectives"] = obj
    np.savez(out, **res)


def _euclid(xy):
    dx = (xy[:, None, 0] - xy[None, :, 0]).astype(np.float32)
    dy = (xy[:, None, 1] - xy[None, :, 1]).astype(np.float32)
    d = np.sqrt((dx * dx).astype(np.float32) + (dy * dy).astype(np.float32)).astype(np.float32)
    np.fill_diagonal(d, 0.0)
    return d


def main():
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(PHILOX_KAT, f, indent=1)
    ref = pyoracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so is not built: run `make -C oracle ref` first")
    knapsack_threshold(ref, os.path.join(HERE, "knapsack_threshold_ref.npz"))
    multitsp(ref, os.path.join(HERE, "multitsp_ref.npz"))
    dominance(ref, os.path.join(HERE, "dominance_ref.npz"))
    loaders(ref, os.path.join(HERE, "loaders_ref.npz"))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
