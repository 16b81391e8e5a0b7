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
     instances, n log-uniform in [50, 2000], P=256, elite/mutants 38.

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
        batch.evolve(args.batch, "launch")
        print(f"batched generations: {batch.member(0).generation} of 512 TSP instances", flush=True)


if __name__ == "__main__":
    main()
