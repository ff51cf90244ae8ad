"""Headline benchmark: DoF/s and fraction of the HBM roofline of the 3D,
degree-4, FP64 DG operator apply on one MI355X.

Workload (the metric's configuration): 64^3 cells on [0,1]^3 with interior
vertices perturbed by up to 10% of h (seed 42), degree 4 on Gauss-Lobatto
nodes (32,768,000 DoF), periodic, divergence-free velocity
(sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy), upwind flux, Gaussian pulse
initial data of width 0.1. One step = one operator apply
dst = M^-1(-L src) over the whole mesh with src/dst resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6


def workload(refine=6, degree=4, dim=3):
    return dict(dim=dim, degree=degree, refine=refine, perturbation=0.1,
                seed=42, velocity_kind=1, flux_alpha=1.0, factor_skew=0.0,
                solution_kind=3, solution_param=[0.5, 0.5, 0.5, 0.1],
                periodic=[1, 1, 1], lo=[0, 0, 0], hi=[1, 1, 1])


def cpu_baseline(threads=None, target_seconds=15.0):
    """The oracle's apply (restated reference algorithm, OpenMP over cells)
    on a bounded sample: 16^3 cells of the same workload."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    from tests.helpers import Cfg
    cores = threads or len(os.sched_getaffinity(0))
    oracle.load().or_set_threads(cores)
    o = Cfg(**workload(refine=4)).oracle_handle()
    u = o.project(0.0)
    o.apply(u, 0.0)  # warm up
    reps, t0 = 0, time.perf_counter()
    while True:
        o.apply(u, 0.0)
        reps += 1
        el = time.perf_counter() - t0
        if el > target_seconds or reps >= 50:
            break
    return dict(value=o.n_dofs * reps / el, unit="DoF/s", cores=cores,
                kind="port",
                sample=f"{reps} applies on 16^3 cells (degree 4, "
                       f"{o.n_dofs} DoF) of the same workload, restated "
                       "reference algorithm (oracle/oracle.cpp), not deal.II")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--refine", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this build has one "
              "MI355X; multi-GPU sharding is described, not run",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config

    cfg = make_config(**workload(refine=args.refine))
    op = AdvectionOperation(cfg)
    src = op.initialize_dof_vector()
    op.project_initial(src, 0.0)
    dst = op.initialize_dof_vector()
    for _ in range(args.warmup):
        op.apply(src, dst, 0.0)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), \
        torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        op.apply(src, dst, 0.0)
    stop.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = start.elapsed_time(stop) / args.steps
    n = op.n_dofs
    dofs_per_s = n / (ms * 1e-3)
    bytes_per_apply = 16.0 * n
    achieved = bytes_per_apply / (ms * 1e-3) / 1e9
    out = {
        "metric": "DoF/s & %HBM peak, 3D p=4 DG apply",
        "value": dofs_per_s,
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D advection, degree 4, 64^3 perturbed "
                               "cells, cellular velocity, FP64 apply",
                   "n_dofs": n, "refine": args.refine},
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
