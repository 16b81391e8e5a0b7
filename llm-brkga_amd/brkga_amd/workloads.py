"""Synthetic problem instances drawn from Philox4x32-10 (no host RNG library).

Instance i of a workload uses key = master seed and counters
(i, stream_tag, j, 0); word w of block j is the w-th output word.  All
arithmetic is FP32/int in a fixed order, so every run on every machine sees
the same instances.  Shapes follow the benchmark configurations:

  * TSP: n cities, coordinates uniform in [0, span)^2.
  * knapsack: weights U{1..100}, values = weight + U{0..10},
    capacity = floor(0.5 * sum(weights)).
  * TSPJ: depot + n cities uniform in [0, 1000)^2, Euclidean travel times,
    job times uniform in [1, 50).
  * bi-objective TSP: two independent coordinate sets of n + 1 points.
"""
from __future__ import annotations

import numpy as np

MASTER_SEED = 0x5EEDBA5E

TAG_COORDS = 0x101
TAG_KNAPSACK = 0x102
TAG_TSPJ_COORDS = 0x103
TAG_TSPJ_JOBS = 0x104
TAG_BITSP_A = 0x105
TAG_BITSP_B = 0x106
TAG_SIZES = 0x107
TAG_MIX = 0x108

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 (Salmon et al., SC'11) over uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32).copy() for x in np.broadcast_arrays(c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + _W0)
            k1 = np.uint32(k1 + _W1)
    return c0, c1, c2, c3


def words(instance: int, tag: int, count_blocks: int, seed: int = MASTER_SEED):
    """[count_blocks, 4] uint32 words for (instance, tag, block j, 0)."""
    j = np.arange(count_blocks, dtype=np.uint32)
    out = philox4x32_10(np.uint32(instance), np.uint32(tag), j, np.uint32(0), seed & 0xFFFFFFFF, seed >> 32)
    return np.stack(out, axis=1)


def unit(u):
    """24 high bits of u as an exact FP32 value in [0, 1)."""
    return (np.asarray(u, np.uint32) >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)


def tsp_coords(instance: int, n: int, span: float = 1000.0, tag: int = TAG_COORDS, seed: int = MASTER_SEED):
    w = words(instance, tag, n, seed)
    xy = np.empty((n, 2), np.float32)
    xy[:, 0] = unit(w[:, 0]) * np.float32(span)
    xy[:, 1] = unit(w[:, 1]) * np.float32(span)
    return xy


def knapsack(instance: int, n: int, seed: int = MASTER_SEED):
    w = words(instance, TAG_KNAPSACK, n, seed)
    weights = (1 + (w[:, 0] % np.uint32(100))).astype(np.int32)
    values = (weights + (w[:, 1] % np.uint32(11)).astype(np.int32)).astype(np.int32)
    capacity = int(weights.astype(np.int64).sum()) // 2
    return weights, values, capacity


def euclid_matrix(xy):
    """FP32 Euclidean matrix with separately rounded products and sum,
    zero diagonal (configs/tsp_config.hpp:359-371)."""
    xy = np.asarray(xy, np.float32)
    dx = (xy[:, None, 0] - xy[None, :, 0]).astype(np.float32)
    dy = (xy[:, None, 1] - xy[None, :, 1]).astype(np.float32)
    s = (dx * dx).astype(np.float32) + (dy * dy).astype(np.float32)
    d = np.sqrt(s.astype(np.float32)).astype(np.float32)
    np.fill_diagonal(d, 0.0)
    return d


def tspj(instance: int, n: int, seed: int = MASTER_SEED):
    xy = tsp_coords(instance, n + 1, 1000.0, TAG_TSPJ_COORDS, seed)
    travel = euclid_matrix(xy)
    w = words(instance, TAG_TSPJ_JOBS, ((n + 1) * (n + 1) + 3) // 4, seed).reshape(-1)[: (n + 1) * (n + 1)]
    jobs = (np.float32(1.0) + unit(w) * np.float32(49.0)).astype(np.float32).reshape(n + 1, n + 1)
    jobs[0, :] = 0.0
    jobs[:, 0] = 0.0
    return travel, jobs


def bitsp_coords(instance: int, n: int, span: float = 1000.0, seed: int = MASTER_SEED):
    return (tsp_coords(instance, n + 1, span, TAG_BITSP_A, seed),
            tsp_coords(instance, n + 1, span, TAG_BITSP_B, seed))


def log_uniform_sizes(count: int, lo: int, hi: int, tag: int = TAG_SIZES, seed: int = MASTER_SEED):
    """count sizes n = round(exp(U[ln lo, ln hi])) from one Philox stream."""
    w = words(0, tag, count, seed)[:, 0]
    u = unit(w).astype(np.float64)
    n = np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    return np.clip(n, lo, hi).astype(np.int32)


TAG_ARRIVALS = 0x109


def request_mix(count: int, seed: int = MASTER_SEED):
    """Mixed serving requests: (kind, n) per request, 60% Euclidean TSP with
    n log-uniform in [20, 5000], 25% knapsack with n uniform in {100..2000},
    15% bi-objective TSP with n uniform in {50..500}."""
    w = words(1, TAG_MIX, count, seed)
    pick = unit(w[:, 0]).astype(np.float64)
    u = unit(w[:, 1]).astype(np.float64)
    out = []
    for i in range(count):
        if pick[i] < 0.60:
            n = int(np.clip(np.rint(np.exp(np.log(20) + u[i] * (np.log(5000) - np.log(20)))), 20, 5000))
            out.append(("tsp", n))
        elif pick[i] < 0.85:
            out.append(("knapsack", 100 + int(w[i, 2] % np.uint32(1901))))
        else:
            out.append(("bitsp", 50 + int(w[i, 2] % np.uint32(451))))
    return out


def poisson_arrivals(count: int, rate: float, seed: int = MASTER_SEED):
    """Arrival times (seconds from 0) of a Poisson stream of `rate` requests
    per second: exponential gaps -ln(1 - u) / rate from one Philox stream."""
    w = words(2, TAG_ARRIVALS, count, seed)[:, 0]
    u = unit(w).astype(np.float64)
    return np.cumsum(-np.log1p(-u) / float(rate))

