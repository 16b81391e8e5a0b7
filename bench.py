#!/usr/bin/env python3
"""Benchmark of the BRKGA generation step on one MI355X.

Workload (the configuration the headline metric is quoted on): Euclidean TSP,
n = 1000 cities with FP32 coordinates uniform in [0, 10000)^2 (Philox-seeded),
population 1024 as 4 islands x 256, elite 38 and mutants 38 per island,
rho = 0.7, ring migration of the top 2 every 50 generations.  A step is one
generation of the whole population; value = generations per second with the
population resident in HBM, replayed from a captured HIP graph.

Also reported: the same steps launched one by one (no graph), the dominant
kernel's roofline fraction (HIP-event timing of each launch, algorithmic
bytes per generation B = 4 L (3P - M) + 12 P), and the CPU oracle timed on a
bounded sample of the same workload.

Batched leg ("batched"): 512 independent Euclidean TSP instances, n
log-uniform in [50, 2000], coordinates uniform in [0, 1000)^2, population
256, elite 38, mutants 38, rho 0.7, 300 generations each.  One solve =
generation 0 + 300 steps.  Timed three ways on the same instances: the
plain per-instance form (one instance after the other, one launch sequence
per generation), the batched launches, and the batched launches replayed
from a captured graph (after the autotuner picked each size bucket's launch
shape).  Every member's best fitness must be identical across the forms.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 50]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "llm-brkga_amd"))

METRIC = "TSP n=1000 P=1024 gens/s; mixed-n batched solves/s + p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0
N_CITIES, POP, ISLANDS, ELITE, MUTANTS, RHO, MIG_INTERVAL, MIG_K = 1000, 1024, 4, 152, 152, 0.7, 50, 2
SEED = 0xB5EED


def algorithmic_bytes_per_generation(L, P, M):
    # decode reads all keys, crossover / elite copy read the selected parent
    # genes, all child rows are written, fitness written, partition reads
    # fitness and writes the rank order
    return 4 * L * (3 * P - M) + 12 * P


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "traffic_k_brkga_step_tsp_n1000_p1024.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(budget_s):
    """The CPU oracle (single thread) on the same instance; generations are
    run until the time budget is spent (at least two)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from brkga_amd import workloads

    xy = workloads.tsp_coords(0, N_CITIES, 10000.0)
    prob = pyoracle.Problem.tsp_coords(xy)
    q = pyoracle.params(POP, ELITE, MUTANTS, RHO, ISLANDS, MIG_INTERVAL, MIG_K, SEED)
    st = pyoracle.BrkgaState(prob, q)
    t0 = time.perf_counter()
    gens = 0
    while gens < 2 or time.perf_counter() - t0 < budget_s:
        st.step()
        gens += 1
    dt = time.perf_counter() - t0
    return {"value": round(gens / dt, 3), "unit": "gens/s", "cores": 1, "kind": "port",
            "sample": f"{gens} generations of the same n=1000 P=1024 4-island workload, CPU oracle (C, -O2, "
                      f"one thread), {dt:.1f} s"}


BATCH_COUNT, BATCH_POP, BATCH_ELITE, BATCH_MUTANTS, BATCH_RHO, BATCH_GENS = 512, 256, 38, 38, 0.7, 300
BATCH_N_LO, BATCH_N_HI = 50, 2000


def batched_leg(brkga_amd, workloads, eng, gens):
    sizes = [int(n) for n in workloads.log_uniform_sizes(BATCH_COUNT, BATCH_N_LO, BATCH_N_HI)]
    probs = [brkga_amd.Problem.tsp_coords(eng, workloads.tsp_coords(1000 + i, n, 1000.0))
             for i, n in enumerate(sizes)]
    params = [brkga_amd.Params(BATCH_POP, BATCH_ELITE, BATCH_MUTANTS, BATCH_RHO, 1, 0, 0, SEED + i)
              for i in range(BATCH_COUNT)]
    batch = brkga_amd.Batch(eng, probs, params)

    def sync():
        return batch.member(0).generation

    def bests():
        return [batch.member(i).best()[0] for i in range(BATCH_COUNT)]

    # warm-up: one batched solve, tuning, and the graph capture
    batch.initialize("launch")
    batch.autotune(reps=2)
    batch.evolve(gens, "graph")
    sync()

    t0 = time.perf_counter()
    batch.initialize("launch")
    batch.evolve(gens, "graph")
    sync()
    t_graph = time.perf_counter() - t0
    ref = bests()

    t0 = time.perf_counter()
    batch.initialize("launch")
    batch.evolve(gens, "launch")
    sync()
    t_launch = time.perf_counter() - t0
    same = bests() == ref

    t0 = time.perf_counter()
    batch.initialize("plain")
    batch.evolve(gens, "plain")
    sync()
    t_plain = time.perf_counter() - t0
    same = same and bests() == ref

    B = sum(algorithmic_bytes_per_generation(n, BATCH_POP, BATCH_MUTANTS) for n in sizes) * gens
    achieved = B / t_graph / 1e9
    out = {
        "workload": f"{BATCH_COUNT} Euclidean TSP instances, n log-uniform in [{BATCH_N_LO},{BATCH_N_HI}] "
                    f"(sum n = {sum(sizes)}), coordinates uniform FP32 in [0,1000)^2, P={BATCH_POP}, "
                    f"elite {BATCH_ELITE}, mutants {BATCH_MUTANTS}, rho {BATCH_RHO}, {gens} generations",
        "solves_per_s": round(BATCH_COUNT / t_graph, 2),
        "solves_per_s_launch": round(BATCH_COUNT / t_launch, 2),
        "solves_per_s_plain": round(BATCH_COUNT / t_plain, 2),
        "speedup_vs_plain": round(t_plain / t_graph, 2),
        "ms_per_batch_generation": round(1000.0 * t_graph / gens, 4),
        "plain_seconds": round(t_plain, 3),
        "results_identical": bool(same),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": B},
        "buckets": batch.buckets(),
    }
    batch.close()
    for p in probs:
        p.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-gens", type=int, default=BATCH_GENS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the island model runs all islands on one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import brkga_amd
    from brkga_amd import workloads

    K, W = args.steps, args.warmup
    eng = brkga_amd.Engine(0)
    xy = workloads.tsp_coords(0, N_CITIES, 10000.0)
    prob = brkga_amd.Problem.tsp_coords(eng, xy)
    prm = brkga_amd.Params(POP, ELITE, MUTANTS, RHO, ISLANDS, MIG_INTERVAL, MIG_K, SEED)
    pop = brkga_amd.Population(prob, prm)
    pop.initialize()

    def sync():
        return pop.generation  # synchronises the stream the steps run on

    # warm-up: plain steps, then one run of the K-step graph (captures it)
    pop.evolve(W)
    pop.evolve(K, graph=True)
    sync()

    t0 = time.perf_counter()
    pop.evolve(K, graph=True)
    sync()
    t_graph = time.perf_counter() - t0

    t0 = time.perf_counter()
    pop.evolve(K)
    sync()
    t_plain = time.perf_counter() - t0

    prof_gens = min(K, 200)
    step_us, close_us = pop.profile(prof_gens)
    best, _, tour = pop.best()
    L, P, M = N_CITIES, POP, MUTANTS
    B = algorithmic_bytes_per_generation(L, P, M)
    achieved = B / (step_us * 1e-6) / 1e9
    gens_per_s = K / t_graph
    out = {
        "metric": METRIC,
        "value": round(gens_per_s, 2),
        "unit": "gens/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W + K,
        "ms_per_step": round(1000.0 * t_graph / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-seeded coordinates, uniform FP32 in [0,10000)^2",
        "config": {"workload": "TSP n=1000, P=1024 = 4 islands x 256, elite 38 + mutants 38 per island, "
                               "rho 0.7, ring migration top-2 every 50 generations",
                   "n": L, "population": P, "islands": ISLANDS, "elite": ELITE, "mutants": M, "rho": RHO,
                   "launch": "hipGraph replay of the K-generation loop"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": "k_brkga_step (generation + decode + fitness)",
                     "kernel_us": round(step_us, 3), "partition_us": round(close_us, 3),
                     "bytes_per_generation": B,
                     "generation_frac": round(B * gens_per_s / (HBM_PEAK_GBS * 1e9), 4)},
        "plain_gens_per_s": round(K / t_plain, 2),
        "best_tour": round(float(best), 3),
        "valid_tour": sorted(tour.tolist()) == list(range(L)),
    }
    if not args.no_batch:
        out["batched"] = batched_leg(brkga_amd, workloads, eng, args.batch_gens)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
