#!/usr/bin/env python3
"""Workload driven under rocprofv3 to profile the generation kernels.

Runs, in order:
  1. a device-to-device copy of --copy-mib MiB through a plain copy kernel
     (torch's elementwise copy), the calibration point for the FETCH_SIZE /
     WRITE_SIZE counters;
  2. `--gens` plain (one launch per kernel) generations of the headline
     workload: TSP n=1000, P=1024 as 4 islands x 256, elite/mutants 38 per
     island, rho 0.7, migration of the top 2 every 50 generations;
  3. with --batch G, G batched generations (one launch per size bucket per
     generation, no graph) of the mixed-size workload: 512 Euclidean TSP
     instances, n log-uniform in [50, 2000], P=256, elite/mutants 38;
  4. with --serve R, R mixed serving requests (60% TSP n log-uniform in
     [20, 5000], 25% greedy knapsack n in [100, 2000], 15% bi-objective TSP
     with NSGA-II population 512; 200-generation deadline) submitted at once
     to the request loop.
With --tune the batched and served parts first autotune their launch shapes
(as the benchmark does).

Example:
  rocprofv3 --kernel-trace --stats --output-format csv -d out -o run -- \
      python3 llm-brkga_amd/profile_workload.py --gens 200
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gens", type=int, default=200)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--serve", type=int, default=0)
    ap.add_argument("--tune", action="store_true", help="autotune launch shapes before the profiled part")
    args = ap.parse_args()

    import torch

    import brkga_amd
    from brkga_amd import workloads

    n_el = args.copy_mib * (1 << 20) // 4
    src = torch.ones(n_el, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy kernel: {args.copy_mib} MiB read + {args.copy_mib} MiB written", flush=True)
    del src, dst

    eng = brkga_amd.Engine(0)
    prob = brkga_amd.Problem.tsp_coords(eng, workloads.tsp_coords(0, args.n, 10000.0))
    pop = brkga_amd.Population(prob, brkga_amd.Params(1024, 152, 152, 0.7, 4, 50, 2, 0xB5EED))
    pop.initialize()
    pop.evolve(args.gens)
    print(f"generations: {pop.generation} of TSP n={args.n} P=1024 (4 islands)", flush=True)
    if args.batch > 0:
        sizes = [int(n) for n in workloads.log_uniform_sizes(512, 50, 2000)]
        probs = [brkga_amd.Problem.tsp_coords(eng, workloads.tsp_coords(1000 + i, n, 1000.0))
                 for i, n in enumerate(sizes)]
        params = [brkga_amd.Params(256, 38, 38, 0.7, 1, 0, 0, 0xB5EED + i) for i in range(512)]
        batch = brkga_amd.Batch(eng, probs, params)
        batch.initialize("launch")
        if args.tune:
            batch.autotune(2)
        batch.evolve(args.batch, "launch")
        print(f"batched generations: {batch.member(0).generation} of 512 TSP instances", flush=True)
    if args.serve > 0:
        srv = brkga_amd.Server(eng, hbm_budget_bytes=8 << 30, ticks_per_round=4, autotune_reps=2 if args.tune else 0)
        ids = []
        for i, (kind, n) in enumerate(workloads.request_mix(args.serve)):
            inst, seed = 5000 + i, 0xB5EED + 0x10000 + i
            if kind == "tsp":
                p = brkga_amd.Problem.tsp_coords(eng, workloads.tsp_coords(inst, n, 1000.0))
            elif kind == "knapsack":
                w, v, cap = workloads.knapsack(inst, n)
                p = brkga_amd.Problem.knapsack(eng, w, v, cap, greedy=True)
            else:
                a, b = workloads.bitsp_coords(inst, n)
                p = brkga_amd.Problem.bitsp_coords(eng, a, b)
            prm = (brkga_amd.Params(512, 0, 0, 0.7, 1, 0, 0, seed) if kind == "bitsp"
                   else brkga_amd.Params(256, 38, 38, 0.7, 1, 0, 0, seed))
            ids.append(srv.submit(p, prm, 200, own_problem=True))
        for r in ids:
            srv.wait(r)
        print(f"served requests: {len(ids)}, ticks {srv.stats()['ticks']}", flush=True)
        srv.stop()


if __name__ == "__main__":
    main()
