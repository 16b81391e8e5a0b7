"""Instance builders shared by the GPU tests: the same data as an engine
problem and as an oracle problem."""
import numpy as np

import brkga_amd
import pyoracle
from brkga_amd import workloads

CONTINUOUS = {"cont_sphere": (pyoracle.SPHERE, -5.12, 5.12), "cont_rastrigin": (pyoracle.RASTRIGIN, -5.12, 5.12),
              "cont_rosenbrock": (pyoracle.ROSENBROCK, -5.0, 10.0)}


def vrp_time_matrix(inst, n):
    """Drive minutes between n + 1 locations (0 = depot): Euclidean over a
    100 x 100 grid with +-20% asymmetric noise."""
    r = np.random.default_rng(1000 + inst)
    xy = r.uniform(0, 100, (n + 1, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    return (d * r.uniform(0.8, 1.2, d.shape)).astype(np.float32)


def cvrp(inst, n):
    """Depot + n customers on a 100 x 100 grid, demands 1..29, capacity 100,
    one vehicle per six customers."""
    r = np.random.default_rng(2000 + inst)
    xy = np.round(r.uniform(0, 100, (n + 1, 2)), 3).astype(np.float32)
    dem = r.integers(1, 30, n + 1).astype(np.int32)
    dem[0] = 0
    return xy, dem, 100, max(3, n // 6)


def make_problem(engine, kind, n, inst=0):
    """(engine problem, oracle problem) over identical data."""
    if kind == "tsp":
        xy = workloads.tsp_coords(inst, n, 10000.0)
        return brkga_amd.Problem.tsp_coords(engine, xy), pyoracle.Problem.tsp_coords(xy)
    if kind == "tsp_matrix":
        d = workloads.euclid_matrix(workloads.tsp_coords(inst, n))
        return brkga_amd.Problem.tsp_matrix(engine, d), pyoracle.Problem.tsp_matrix(d)
    if kind in ("knap_greedy", "knap_threshold"):
        w, v, cap = workloads.knapsack(inst, n)
        g = kind == "knap_greedy"
        return brkga_amd.Problem.knapsack(engine, w, v, cap, g), pyoracle.Problem.knapsack(w, v, cap, g)
    if kind == "tspj":
        t, j = workloads.tspj(inst, n)
        return brkga_amd.Problem.tspj(engine, t, j), pyoracle.Problem.tspj(t, j)
    if kind == "bitsp":
        a, b = workloads.bitsp_coords(inst, n)
        return brkga_amd.Problem.bitsp_coords(engine, a, b), pyoracle.Problem.bitsp_coords(a, b)
    if kind == "bitsp_matrix":
        a, b = workloads.bitsp_coords(inst, n)
        ma, mb = workloads.euclid_matrix(a), workloads.euclid_matrix(b)
        return brkga_amd.Problem.bitsp_matrix(engine, ma, mb), pyoracle.Problem.bitsp_matrix(ma, mb)
    if kind in CONTINUOUS:
        fn, lo, hi = CONTINUOUS[kind]
        return brkga_amd.Problem.continuous(engine, fn, n, lo, hi), pyoracle.Problem.continuous(fn, n, lo, hi)
    if kind == "jobshop":
        r = np.random.default_rng(3000 + inst)
        t = r.uniform(1, 50, n).astype(np.float32)
        M = int(r.integers(2, 33))
        return brkga_amd.Problem.job_scheduling(engine, t, M), pyoracle.Problem.job_scheduling(t, M)
    if kind == "zdt1":
        return brkga_amd.Problem.zdt1(engine, n), pyoracle.Problem.zdt1(n)
    if kind in ("vrp_days", "vrp_days_tight"):
        tm = vrp_time_matrix(inst, n)
        stops, minutes = (12, 480.0) if kind == "vrp_days" else (4, 150.0)
        return (brkga_amd.Problem.vrp_days(engine, tm, stops, minutes),
                pyoracle.Problem.vrp_days(tm, stops, minutes))
    if kind == "movrp":
        xy, dem, cap, veh = cvrp(inst, n)
        return brkga_amd.Problem.movrp(engine, xy, dem, cap, veh), pyoracle.Problem.movrp(xy, dem, cap, veh)
    raise ValueError(kind)
