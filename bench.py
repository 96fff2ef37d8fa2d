#!/usr/bin/env python3
"""Benchmark: BLR-LU factor + refined solve at N=32768 on one MI355X.

Workload (the metric's configuration): 1-D Laplace kernel
A_ij = 1/(|x_i - x_j| + 1e-3) + 1e-2 delta_ij, x = sorted U[0,1) points from
std::mt19937(0) (src/util/experiment_setup.cpp:12-23), BLR 64 x 64 grid of
512 x 512 tiles, weak admissibility, eps = 1e-10, 8 right-hand sides ~ N(0,1),
FP64 iterative refinement with the exact matrix to backward error <= 1e-10.

A step = restore the compressed matrix (one device copy) + getrf + refined
solve.  Construction is timed separately (as examples/h_lu.cpp:24-43 does).

Two paths are timed: plain FP64 (eps 1e-10) and mixed precision (FP32 tile
storage and recompression at eps 1e-6, FP64 diagonal getrf, FP64 refinement
against the uncompressed kernel).  `value` is the mixed-precision time.

The `roofline` object covers one kernel class of the headline path (its
launches are bracketed by HIP events on the library's stream during the timed
steps; flops and bytes come from the library's op log).  `cpu_baseline` times
the CPU restatement of the reference algorithm (oracle/) on the first
elimination steps of the same FP64 matrix and scales by the tile-op count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
PEAK_HBM_GBS = 8000.0
RIDGE = {"f64": 78.6e12 / 8e12, "f32": 157.3e12 / 8e12}


def cpu_baseline(A64, nb: int, steps: int, nrhs: int):
    """Oracle BLR-LU (rounded addition, FP64, one thread) on the first `steps`
    elimination steps of the downloaded FP64 matrix, scaled to all steps by the
    number of tile operations."""
    from threadpoolctl import threadpool_limits

    from oracle import frank_oracle as O

    t_dl = time.perf_counter()
    blocks, rr, cr = [], [], []
    r0 = 0
    for i in range(nb):
        rr.append((r0, A64.tile_info(i, 0)[2]))
        r0 += rr[-1][1]
    c0 = 0
    for j in range(nb):
        cr.append((c0, A64.tile_info(0, j)[3]))
        c0 += cr[-1][1]
    for i in range(nb):
        for j in range(nb):
            t = A64.tile(i, j)
            blocks.append(O.LowRank(t[0], t[1], t[2], 1e-10) if isinstance(t, tuple) else t)
    H = O.Hier(nb, nb, blocks, rr, cr)
    t_dl = time.perf_counter() - t_dl
    with threadpool_limits(1):
        t0 = time.perf_counter()
        O.blr_getrf(H, lra="rounded_addition", max_steps=steps)
        dt = time.perf_counter() - t0
    upd = [(nb - 1 - k) ** 2 + 2 * (nb - 1 - k) + 1 for k in range(nb)]  # updates + trsm + getrf per step
    frac = sum(upd[:steps]) / sum(upd)
    print(f"[bench] cpu baseline: {steps} steps in {dt:.1f}s (download {t_dl:.1f}s), fraction of work {frac:.3f}",
          file=sys.stderr, flush=True)
    return {"value": dt / frac, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle/frank_oracle.py blr_getrf (FP64, rounded addition, 1 BLAS thread) on the first "
                      f"{steps} of {nb} elimination steps of the same FP64 BLR matrix (downloaded from the GPU), "
                      f"{dt:.1f}s measured, scaled by the per-step tile-op count ({frac:.3f} of the total); "
                      f"factorization only, the {nrhs}-RHS solve adds under 1%"}


def measured_traffic(cls: str):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated against a
    copy of known size), or None when no summary exists for the class."""
    path = os.path.join(ROOT, "profiles", f"{cls}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    if "hbm_bytes_per_launch" not in vals:
        return None
    return vals["hbm_bytes_per_launch"]


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
    ap.add_argument("--roofline-class", default="getrf_tile",
                    help="kernel class whose launches are event-timed for the roofline object")
    ap.add_argument("--cpu-steps", type=int, default=3, help="elimination steps in the CPU baseline sample (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (calibrates HBM byte counters under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(the sharding design is described in the documentation)", file=sys.stderr)
        sys.exit(2)

    import torch

    import frank_amd as fa
    from frank_amd.inputs import normal_rhs, sorted_uniform_points

    print(f"[bench] N={args.n} nb={args.nb} eps={args.eps} nrhs={args.nrhs}", file=sys.stderr, flush=True)
    if args.copy_probe:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ctx = fa.default_context()
    pts = [sorted_uniform_points(args.n)]
    b = torch.from_numpy(normal_rhs(args.n, args.nrhs)).cuda()
    head = "mixed" if args.mode in ("both", "mixed") else "f64"

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
        ctx.time_class(args.roofline_class if name == head else None)
        ctx.reset_op_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = []
        for s_ in range(args.steps):
            res.append(step())
            print(f"[bench] {name} step {s_}: iters {res[-1][0]} berr {res[-1][1]:.3e}", file=sys.stderr, flush=True)
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ops = ctx.op_stats()
        ctx.time_class(None)
        cnt = ctx.counters(reset=True)
        print(f"[bench] {name}: {dt * 1e3:.1f} ms/step; per step: {cnt['svd_problems'] / args.steps:.0f} small SVDs, "
              f"{cnt['svd_sweeps'] / max(cnt['svd_problems'], 1):.2f} Jacobi sweeps each (max {cnt['svd_max_sweeps']})",
              file=sys.stderr, flush=True)
        for k, v in sorted(ops.items()):
            extra = f", {v['ms'] / args.steps:.2f} ms/step event-timed" if v["timed_launches"] else ""
            print(f"[bench]   {k:12s} {v['launches'] / args.steps:8.0f} launches/step "
                  f"{v['flops'] / args.steps / 1e9:9.2f} GFLOP/step {v['bytes'] / args.steps / 1e9:8.3f} GB/step{extra}",
                  file=sys.stderr, flush=True)
        return {"s": dt, "berr": max(r[1] for r in res), "iters": res[-1][0], "ops": ops}

    out_modes, cpu = {}, None
    if args.mode in ("both", "f64"):
        A64, tb64 = build(args.eps)
        out_modes["f64"] = run("f64", A64)
        out_modes["f64"]["build_s"] = tb64
        if args.cpu_steps > 0:
            cpu = cpu_baseline(A64, args.nb, args.cpu_steps, args.nrhs)
        del A64
    if args.mode in ("both", "mixed"):
        A64m, tbm = build(args.eps_mixed)
        A32 = A64m.convert("f32")
        del A64m
        out_modes["mixed"] = run("mixed", A32)
        out_modes["mixed"]["build_s"] = tbm
    h = out_modes[head]
    dt = h["s"]
    steps = args.steps
    alg_flops = sum(v["flops"] for v in h["ops"].values()) / steps
    roof = None
    rc = h["ops"].get(args.roofline_class)
    if rc and rc["timed_launches"] > 0 and rc["ms"] > 0:
        prec = "f64" if (head == "f64" or args.roofline_class.startswith("getrf")) else "f32"
        secs = rc["ms"] / 1e3
        intensity = rc["flops"] / max(rc["bytes"], 1.0)
        if intensity < RIDGE[prec]:
            ach = rc["bytes"] / secs / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS}
        else:
            ach = rc["flops"] / secs / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                    "frac": ach / PEAK_TFLOPS[prec]}
        traffic = measured_traffic(args.roofline_class)
        roof.update({"traffic": traffic, "kernel_class": args.roofline_class, "precision": prec,
                     "launches_per_step": rc["launches"] / steps,
                     "avg_launch_us": rc["ms"] * 1e3 / rc["timed_launches"],
                     "alg_flops_per_launch": rc["flops"] / rc["launches"],
                     "alg_bytes_per_launch": rc["bytes"] / rc["launches"]})
    n = args.n
    out = {
        "metric": "BLR-LU solve time & TFLOP/s @N=32768, err<=1e-10: mixed-prec vs FP64 GPU",
        "value": dt,
        "unit": "s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"BLR-LU factor + refined solve, 1-D Laplace + 1e-2 I, N={n}, "
                               f"{args.nb}x{args.nb} tiles of {n // args.nb}, eps={args.eps}, {args.nrhs} RHS, "
                               f"backward error <= {args.tol}",
                   "mixed_eps": args.eps_mixed if head == "mixed" else None},
        "berr": h["berr"],
        "refine_iters": h["iters"],
        "construction_s": h["build_s"],
        "alg_tflops": alg_flops / dt / 1e12,
        "effective_tflops": (2.0 / 3.0) * n ** 3 / dt / 1e12,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "f64" in out_modes:
        out["f64_s"] = out_modes["f64"]["s"]
        out["f64_berr"] = out_modes["f64"]["berr"]
        out["f64_alg_tflops"] = sum(v["flops"] for v in out_modes["f64"]["ops"].values()) / steps / out["f64_s"] / 1e12
    if "f64" in out_modes and "mixed" in out_modes:
        out["speedup_vs_f64"] = out_modes["f64"]["s"] / out_modes["mixed"]["s"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
