#!/usr/bin/env python3
"""Benchmark: BLR-LU factor + refined solve at N=32768 on one MI355X.

Workload (the metric's configuration): 1-D Laplace kernel
A_ij = 1/(|x_i - x_j| + 1e-3) + 1e-2 delta_ij, x = sorted U[0,1) points from
std::mt19937(0) (src/util/experiment_setup.cpp:12-23), BLR 64 x 64 grid of
512 x 512 tiles, weak admissibility, eps = 1e-10, 8 right-hand sides ~ N(0,1),
FP64 iterative refinement with the exact matrix to backward error <= 1e-10.

A step = restore the compressed matrix (one device copy) + getrf + refined
solve.  Construction is timed separately (as examples/h_lu.cpp:24-43 does).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=64)
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--mode", choices=["f64"], default="f64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import frank_amd as fa
    from frank_amd.inputs import normal_rhs, sorted_uniform_points

    print(f"[bench] N={args.n} nb={args.nb} eps={args.eps} nrhs={args.nrhs}", file=sys.stderr, flush=True)
    ctx = fa.default_context()
    pts = [sorted_uniform_points(args.n)]
    t0 = time.perf_counter()
    A = fa.BlrMatrix.from_kernel(pts, args.nb, eps=args.eps, shift=1e-2, kcap=64, ctx=ctx)
    ctx.synchronize()
    t_build = time.perf_counter() - t0
    ranks = A.ranks()
    print(f"[bench] construction {t_build:.2f}s, max rank {ranks.max()}, mean rank {ranks[ranks>=0].mean():.2f}",
          file=sys.stderr, flush=True)
    A.set_lra("rounded_addition")
    F = A.clone()
    b = torch.from_numpy(normal_rhs(args.n, args.nrhs)).cuda()
    x = torch.zeros_like(b)

    def step():
        F.copy_from(A)
        F.getrf()
        return F.refine(b, x, tol=args.tol)

    for w in range(args.warmup):
        it, be = step()
        print(f"[bench] warmup {w}: iters {it} berr {be:.3e}", file=sys.stderr, flush=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = []
    for s in range(args.steps):
        res.append(step())
        print(f"[bench] step {s}: iters {res[-1][0]} berr {res[-1][1]:.3e}", file=sys.stderr, flush=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    out = {
        "metric": "BLR-LU solve time & TFLOP/s @N=32768, err<=1e-10: mixed-prec vs FP64 GPU",
        "value": dt,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"BLR-LU refined solve, 1-D Laplace + 1e-2 I, N={args.n}, {args.nb}x{args.nb} "
                               f"tiles, eps={args.eps}, {args.nrhs} RHS, backward error <= {args.tol}"},
        "berr": max(r[1] for r in res),
        "refine_iters": res[-1][0],
        "construction_s": t_build,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
