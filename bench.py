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
    ap.add_argument("--mode", choices=["both", "f64", "mixed"], default="both")
    ap.add_argument("--eps-mixed", type=float, default=1e-6,
                    help="compression / recompression accuracy of the FP32-storage factorization")
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
    b = torch.from_numpy(normal_rhs(args.n, args.nrhs)).cuda()

    def build(eps):
        t0 = time.perf_counter()
        A = fa.BlrMatrix.from_kernel(pts, args.nb, eps=eps, shift=1e-2, kcap=64, ctx=ctx)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        A.set_lra("rounded_addition")
        r = A.ranks()
        print(f"[bench] construction eps={eps}: {dt:.2f}s, max rank {r.max()}, mean {r[r >= 0].mean():.2f}",
              file=sys.stderr, flush=True)
        return A, dt

    def run(name, A):
        F = A.clone()
        x = torch.zeros_like(b)

        def step():
            F.copy_from(A)
            F.getrf()
            return F.refine(b, x, tol=args.tol, max_iter=60)

        for w in range(args.warmup):
            it, be = step()
            print(f"[bench] {name} warmup {w}: iters {it} berr {be:.3e}", file=sys.stderr, flush=True)
        ctx.synchronize()
        torch.cuda.synchronize()
        ctx.counters(reset=True)
        t0 = time.perf_counter()
        res = []
        for s_ in range(args.steps):
            res.append(step())
            print(f"[bench] {name} step {s_}: iters {res[-1][0]} berr {res[-1][1]:.3e}", file=sys.stderr, flush=True)
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        cnt = ctx.counters(reset=True)
        print(f"[bench] {name}: {dt * 1e3:.1f} ms/step; per step: {cnt['svd_problems'] / args.steps:.0f} small SVDs, "
              f"{cnt['svd_sweeps'] / max(cnt['svd_problems'], 1):.2f} Jacobi sweeps each (max {cnt['svd_max_sweeps']})",
              file=sys.stderr, flush=True)
        return dt, max(r[1] for r in res), res[-1][0]

    out_modes = {}
    if args.mode in ("both", "f64"):
        A64, tb64 = build(args.eps)
        out_modes["f64"] = run("f64", A64) + (tb64,)
        del A64
    if args.mode in ("both", "mixed"):
        A64m, tbm = build(args.eps_mixed)
        A32 = A64m.convert("f32")
        del A64m
        out_modes["mixed"] = run("mixed", A32) + (tbm,)
    head = "mixed" if "mixed" in out_modes else "f64"
    dt, berr, iters, tb = out_modes[head]
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
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"BLR-LU factor + refined solve, 1-D Laplace + 1e-2 I, N={args.n}, "
                               f"{args.nb}x{args.nb} tiles of {args.n // args.nb}, eps={args.eps}, {args.nrhs} RHS, "
                               f"backward error <= {args.tol}"},
        "berr": berr,
        "refine_iters": iters,
        "construction_s": tb,
    }
    if "f64" in out_modes:
        out["f64_s"] = out_modes["f64"][0]
        out["f64_berr"] = out_modes["f64"][1]
    if "f64" in out_modes and "mixed" in out_modes:
        out["speedup_vs_f64"] = out_modes["f64"][0] / out_modes["mixed"][0]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
