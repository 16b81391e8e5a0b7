"""Instance builders shared by the GPU tests: the same data as an engine
problem and as an oracle problem."""
import brkga_amd
import pyoracle
from brkga_amd import workloads


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
    raise ValueError(kind)
