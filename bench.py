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
shape); the batched figure is the faster of the two batched modes.  Every
member's best fitness must be identical across the forms.

Served leg ("served"): Poisson arrival streams (rates 50 ... 2000 req/s, and
all requests at once) of mixed requests: 60% Euclidean TSP with n
log-uniform in [20, 5000], 25% greedy knapsack with n uniform in
{100..2000}, 15% bi-objective TSP (NSGA-II, population 512) with n uniform
in {50..500}; TSP and knapsack populations 256 with elite/mutants 38; every
request has a 200-generation deadline.  The same trace is served by the
request loop (HBM-budget admission, continuous batching) and by the plain
per-instance form (one request after the other, plain launches); both report
solves/s and p50/p99 latency (arrival to result on the host, problem upload
included), and every request's best fitness must agree between the two.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 50]
"""
import argparse
import ctypes
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "llm-brkga_amd"))

METRIC = "TSP n=1000 P=1024 gens/s; mixed-n batched solves/s + p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0
N_CITIES, POP, ISLANDS, ELITE, MUTANTS, RHO, MIG_INTERVAL, MIG_K = 1000, 1024, 4, 152, 152, 0.7, 50, 2
SEED = 0xB5EED


class HipClock:
    """Device-side span of enqueued work: HIP events recorded on the null
    stream (the stream every timed call here runs on) through the HIP runtime
    the engine library already loaded, plus a device-wide synchronisation."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ev = [ctypes.c_void_p(), ctypes.c_void_p()]
        for e in self.ev:
            self._ok(self.hip.hipEventCreate(ctypes.byref(e)), "hipEventCreate")

    @staticmethod
    def _ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {rc}")

    def sync(self):
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def mark(self, i):
        self._ok(self.hip.hipEventRecord(self.ev[i], None), "hipEventRecord")

    def span_ms(self):
        ms = ctypes.c_float()
        self._ok(self.hip.hipEventElapsedTime(ctypes.byref(ms), self.ev[0], self.ev[1]), "hipEventElapsedTime")
        return ms.value


def algorithmic_bytes_per_generation(L, P, M):
    # decode reads all keys, crossover / elite copy read the selected parent
    # genes, all child rows are written, fitness written, partition reads
    # fitness and writes the rank order
    return 4 * L * (3 * P - M) + 12 * P


def traffic_from_profiles():
    """HBM bytes per k_step_one launch (FETCH_SIZE + WRITE_SIZE, calibrated
    against a copy kernel) from the committed profile summary."""
    path = os.path.join(ROOT, "profiles", "k_step_one_tsp_n1000_p1024.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def cpu_identity():
    """(nproc, CPU model name) of this host."""
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    if not model and os.path.exists("/proc/cpuinfo"):
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    return os.cpu_count() or 1, model


def reference_loop(xy, n, pop, islands, elite, mutants, gens, mig_interval, mig_k, threads, runs):
    """The CPU restatement of the reference's host loop (oracle/cpu_reference_loop:
    Individual objects, n x n matrix, std::sort decode, mt19937 crossover, sort
    of the members every generation): one warm-up run, then `runs` timed runs;
    the median."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "cpu_reference_loop")
    with tempfile.NamedTemporaryFile(suffix=".f32", delete=False) as f:
        np.ascontiguousarray(xy, np.float32).tofile(f)
        path = f.name
    try:
        out = subprocess.run([exe, path, str(n), str(pop), str(islands), str(elite), str(mutants), str(RHO),
                              str(gens), str(mig_interval), str(mig_k), str(threads), str(runs), "12345"],
                             capture_output=True, text=True, timeout=600, check=True).stdout
    finally:
        os.unlink(path)
    return json.loads(out)


def cpu_baseline(budget_s):
    """CPU baselines on this host's cores: the restatement of the reference's
    host loop for the headline workload (4 islands x 256, n = 1000) and for
    TSP n = 52, P = 256, 500 generations, single-threaded (as the reference's
    evaluate_all_cpu runs) and with OpenMP over the members on every core of
    this host's share; plus the flat C oracle (single thread) as a secondary
    figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from brkga_amd import workloads

    nproc, model = cpu_identity()
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = nproc
    threads = max(1, min(16, share))
    xy2 = workloads.tsp_coords(0, N_CITIES, 10000.0)
    xy1 = workloads.tsp_coords(52, 52, 1000.0)
    Pi, Ei, Mi = POP // ISLANDS, ELITE // ISLANDS, MUTANTS // ISLANDS
    c2_gens = 20
    c2_1 = reference_loop(xy2, N_CITIES, Pi, ISLANDS, Ei, Mi, c2_gens, MIG_INTERVAL, MIG_K, 1, 5)
    c2_t = reference_loop(xy2, N_CITIES, Pi, ISLANDS, Ei, Mi, c2_gens, MIG_INTERVAL, MIG_K, threads, 5)
    c1_1 = reference_loop(xy1, 52, 256, 1, 51, 25, 500, 0, 0, 1, 5)
    c1_t = reference_loop(xy1, 52, 256, 1, 51, 25, 500, 0, 0, threads, 5)

    # batched instance set: every 32nd of the 512 instances (sizes across the
    # whole distribution), 30 generations each; solves/s scaled to 300
    sizes = [int(v) for v in workloads.log_uniform_sizes(BATCH_COUNT, BATCH_N_LO, BATCH_N_HI)]
    sample_idx = list(range(0, BATCH_COUNT, 32))
    per_solve = []
    for i in sample_idx:
        r = reference_loop(workloads.tsp_coords(1000 + i, sizes[i], 1000.0), sizes[i], BATCH_POP, 1, BATCH_ELITE,
                           BATCH_MUTANTS, 30, 0, 0, threads, 1)
        per_solve.append(r["median_s"] * BATCH_GENS / 30.0)
    batch_cpu = {"solves_per_s": round(len(per_solve) / sum(per_solve), 3), "threads": threads,
                 "sample": f"instances {sample_idx[0]}, {sample_idx[1]}, ... {sample_idx[-1]} of the batched set "
                           f"(n = {min(sizes[i] for i in sample_idx)}..{max(sizes[i] for i in sample_idx)}), "
                           f"30 generations each, scaled to {BATCH_GENS}"}

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
    pick = lambda r: {"gens_per_s": round(r["gens_per_s"], 3), "median_s": round(r["median_s"], 4),
                      "threads": r["threads"], "runs_s": [round(x, 4) for x in r["runs_s"]]}
    return {"value": round(c2_t["gens_per_s"], 3), "unit": "gens/s", "cores": threads, "kind": "port",
            "sample": f"{c2_gens} generations of the headline workload (4 islands x 256, n = 1000, migration "
                      f"top-2 every 50) per run, median of 5 runs after a warm-up, in the CPU restatement of the "
                      f"reference's host loop (Individual objects, n x n FP32 matrix, std::sort decode, mt19937 "
                      f"crossover, member sort per generation) with OpenMP over the members on {threads} threads",
            "nproc": nproc, "cpu_model": model,
            "headline_single_thread": pick(c2_1), "headline_all_threads": pick(c2_t),
            "tsp52_p256_500_generations": {"single_thread": pick(c1_1), "all_threads": pick(c1_t)},
            "batched_set": batch_cpu,
            "oracle_single_thread": {"gens_per_s": round(gens / dt, 3),
                                     "sample": f"{gens} generations of the headline workload, flat-array C oracle "
                                               f"(-O2, one thread), {dt:.1f} s"}}


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

    # the batched form's value: the faster of its two launch modes (both run
    # the same kernels and give the same results)
    t_best, mode = (t_graph, "graph") if t_graph <= t_launch else (t_launch, "launch")
    B = sum(algorithmic_bytes_per_generation(n, BATCH_POP, BATCH_MUTANTS) for n in sizes) * gens
    achieved = B / t_best / 1e9
    out = {
        "workload": f"{BATCH_COUNT} Euclidean TSP instances, n log-uniform in [{BATCH_N_LO},{BATCH_N_HI}] "
                    f"(sum n = {sum(sizes)}), coordinates uniform FP32 in [0,1000)^2, P={BATCH_POP}, "
                    f"elite {BATCH_ELITE}, mutants {BATCH_MUTANTS}, rho {BATCH_RHO}, {gens} generations",
        "solves_per_s": round(BATCH_COUNT / t_best, 2),
        "mode": mode,
        "solves_per_s_graph": round(BATCH_COUNT / t_graph, 2),
        "solves_per_s_launch": round(BATCH_COUNT / t_launch, 2),
        "solves_per_s_plain": round(BATCH_COUNT / t_plain, 2),
        "speedup_vs_plain": round(t_plain / t_best, 2),
        "ms_per_batch_generation": round(1000.0 * t_best / gens, 4),
        "plain_seconds": round(t_plain, 3),
        "results_identical": bool(same),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": B},
        "buckets": batch.buckets(),
        "lanes": batch.lanes(),
    }
    batch.close()
    for p in probs:
        p.close()
    return out


SERVE_RATES = (50, 100, 200, 500, 1000, 2000, 0)  # 0: every request arrives at t = 0
SERVE_GENS, SERVE_POP, SERVE_ELITE, SERVE_MUTANTS, SERVE_NSGA_POP = 200, 256, 38, 38, 512


def serve_requests(brkga_amd, workloads, count):
    reqs = []
    for i, (kind, n) in enumerate(workloads.request_mix(count)):
        inst, seed = 5000 + i, SEED + 0x10000 + i
        if kind == "tsp":
            data = workloads.tsp_coords(inst, n, 1000.0)
        elif kind == "knapsack":
            data = workloads.knapsack(inst, n)
        else:
            data = workloads.bitsp_coords(inst, n)
        if kind == "bitsp":
            prm = brkga_amd.Params(SERVE_NSGA_POP, 0, 0, 0.7, 1, 0, 0, seed)
        else:
            prm = brkga_amd.Params(SERVE_POP, SERVE_ELITE, SERVE_MUTANTS, 0.7, 1, 0, 0, seed)
        reqs.append((kind, n, data, prm))
    return reqs


def serve_problem(brkga_amd, eng, kind, data):
    if kind == "tsp":
        return brkga_amd.Problem.tsp_coords(eng, data)
    if kind == "knapsack":
        w, v, cap = data
        return brkga_amd.Problem.knapsack(eng, w, v, cap, greedy=True)
    a, b = data
    return brkga_amd.Problem.bitsp_coords(eng, a, b)


def percentiles(lat):
    a = sorted(lat)
    pick = lambda q: a[min(len(a) - 1, int(round(q * (len(a) - 1))))]
    return round(1000.0 * pick(0.5), 3), round(1000.0 * pick(0.99), 3)


def served_run(brkga_amd, eng, reqs, arrivals, budget, tune=0):
    srv = brkga_amd.Server(eng, hbm_budget_bytes=budget, ticks_per_round=4, autotune_reps=tune)
    ids, t_sub = [], []
    t0 = time.perf_counter()
    for (kind, _, data, prm), ta in zip(reqs, arrivals):
        wait = ta - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        p = serve_problem(brkga_amd, eng, kind, data)
        ids.append(srv.submit(p, prm, SERVE_GENS, own_problem=True))
        t_sub.append(time.perf_counter() - t0)
    res = [srv.wait(r) for r in ids]
    stats = srv.stats()
    srv.stop()
    done = [ts + (info["done_us"] - info["submit_us"]) * 1e-6 for ts, (_, _, info) in zip(t_sub, res)]
    lat = [d - ta for d, ta in zip(done, arrivals)]
    span = max(done) - float(arrivals[0])
    return [r[0] for r in res], lat, span, stats


def plain_run(brkga_amd, eng, reqs, arrivals):
    fits, lat, done = [], [], []
    t0 = time.perf_counter()
    for (kind, _, data, prm), ta in zip(reqs, arrivals):
        wait = ta - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        p = serve_problem(brkga_amd, eng, kind, data)
        pop = brkga_amd.Population(p, prm)
        pop.initialize()
        pop.evolve(SERVE_GENS)
        fits.append(pop.best()[0])
        pop.close()
        p.close()
        t = time.perf_counter() - t0
        done.append(t)
        lat.append(t - ta)
    return fits, lat, max(done) - float(arrivals[0])


def served_leg(brkga_amd, workloads, eng, count, rates, budget):
    reqs = serve_requests(brkga_amd, workloads, count)
    # warm-up: the whole mix at once, tuning the launch shape of every
    # (kind, size class) it meets; the engine keeps the shapes
    served_run(brkga_amd, eng, reqs, [0.0] * count, budget, tune=2)
    rows, all_same = [], True
    for rate in rates:
        arrivals = workloads.poisson_arrivals(count, rate) if rate > 0 else [0.0] * count
        s_fit, s_lat, s_span, st = served_run(brkga_amd, eng, reqs, arrivals, budget)
        p_fit, p_lat, p_span = plain_run(brkga_amd, eng, reqs, arrivals)
        same = all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(s_fit, p_fit))
        all_same = all_same and same
        s50, s99 = percentiles(s_lat)
        p50, p99 = percentiles(p_lat)
        rows.append({"rate_req_per_s": rate if rate > 0 else "all at t=0", "requests": count,
                     "served": {"solves_per_s": round(count / s_span, 2), "p50_ms": s50, "p99_ms": s99,
                                "peak_resident_bytes": st["peak_bytes"], "ticks": st["ticks"], "lanes": st["lanes"],
                                "rounds": st["rounds"], "worker_wait_ms": round(st["wait_us"] / 1000, 2),
                                "worker_host_ms": round(st["host_us"] / 1000, 2)},
                     "plain": {"solves_per_s": round(count / p_span, 2), "p50_ms": p50, "p99_ms": p99},
                     "speedup_solves_per_s": round(p_span / s_span, 2), "results_identical": bool(same)})
    return {"workload": f"Poisson arrivals of {count} mixed requests per rate: 60% TSP n log-uniform [20,5000], "
                        f"25% greedy knapsack n uniform [100,2000], 15% bi-objective TSP (NSGA-II, N=512) "
                        f"n uniform [50,500]; P=256 elite/mutants 38 otherwise; {SERVE_GENS}-generation deadline",
            "hbm_budget_bytes": budget, "rates": rows, "results_identical": bool(all_same)}


def step_items(n, block):
    """Keys per thread of the step kernel instantiation that runs n keys."""
    return 4 if n <= 4 * block else 8 if n <= 8 * block else 16


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=4.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-gens", type=int, default=BATCH_GENS)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", type=str, default=",".join(str(r) for r in SERVE_RATES))
    ap.add_argument("--serve-budget-gib", type=float, default=8.0)
    ap.add_argument("--step-block", type=int, default=0,
                    help="step workgroup size (64..1024); 0 = autotune it (under a profiler, pin it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the island model runs all islands on one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import brkga_amd
    from brkga_amd import workloads

    K, W = args.steps, args.warmup
    eng = brkga_amd.Engine(0)
    # the HIP runtime and the events are set up first, so that nothing but a
    # synchronisation separates the last untimed steps from the timed ones
    clock = HipClock()
    xy = workloads.tsp_coords(0, N_CITIES, 10000.0)
    prob = brkga_amd.Problem.tsp_coords(eng, xy)
    prm = brkga_amd.Params(POP, ELITE, MUTANTS, RHO, ISLANDS, MIG_INTERVAL, MIG_K, SEED)
    pop = brkga_amd.Population(prob, prm)
    pop.initialize()

    def sync():
        return pop.generation  # synchronises the stream the steps run on

    # warm-up: W plain steps, then the workgroup-size tuning (its timed
    # launches rewrite the same rows, so the population does not advance)
    # and the capture of the K-step graph, which is replayed once untimed
    pop.evolve(W)
    tune_reps = 0
    if args.step_block:
        pop.set_block(args.step_block)
        block = args.step_block
    else:
        tune_reps = 20
        block = pop.autotune(tune_reps)
    pop.evolve(K, graph=True)
    clock.sync()
    t0 = time.perf_counter()
    pop.evolve(K, graph=True)
    clock.sync()
    t_graph = time.perf_counter() - t0

    # device span of the same K-step graph, in a replay of its own (events
    # on the stream around it), after the timed region
    clock.sync()
    clock.mark(0)
    pop.evolve(K, graph=True)
    clock.mark(1)
    clock.sync()
    device_ms = clock.span_ms()

    clock.sync()
    t0 = time.perf_counter()
    pop.evolve(K)
    clock.sync()
    t_plain = time.perf_counter() - t0

    step_us, close_us = pop.profile(200)
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
        "warmup": W,
        "ms_per_step": round(1000.0 * t_graph / K, 5),
        "device_ms_per_step": round(device_ms / K, 5),
        "device_gens_per_s": round(K / (device_ms * 1e-3), 2),
        "untimed_before_timing": {"plain_steps": W, "autotune_timed_launches_per_size": tune_reps,
                                  "graph_capture_replay_steps": K},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-seeded coordinates, uniform FP32 in [0,10000)^2",
        "config": {"workload": "TSP n=1000, P=1024 = 4 islands x 256, elite 38 + mutants 38 per island, "
                               "rho 0.7, ring migration top-2 every 50 generations",
                   "n": L, "population": P, "islands": ISLANDS, "elite": ELITE, "mutants": M, "rho": RHO,
                   "launch": "hipGraph replay of the K-generation loop", "step_block": block},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": f"brkga::k_step_one<0, {block}, {step_items(L, block)}> "
                               "(island rank + crossover / mutants / elite copy + decode + fitness)",
                     "kernel_us": round(step_us, 3), "partition_us": round(close_us, 3),
                     "bytes_per_generation": B,
                     "generation_frac": round(B * gens_per_s / (HBM_PEAK_GBS * 1e9), 4)},
        "plain_gens_per_s": round(K / t_plain, 2),
        "best_tour": round(float(best), 3),
        "valid_tour": sorted(tour.tolist()) == list(range(L)),
    }
    if not args.no_batch:
        out["batched"] = batched_leg(brkga_amd, workloads, eng, args.batch_gens)
    if not args.no_serve:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        out["served"] = served_leg(brkga_amd, workloads, eng, args.serve_requests, rates,
                                   int(args.serve_budget_gib * (1 << 30)))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
