"""Benchmark: solve TFLOP/s at FP64 backward error, mixed-precision IR LU vs
plain FP64 LU, on one MI355X.

Default workload: one dense FP64 system, n = 16384, 1 right-hand side,
A = U diag(sigma) V^T with U, V the Q factors of seeded N(0,1) matrices and
sigma log-spaced over [1, 1e4] (kappa_2 = 1e4), b = A x_true with x_true
N(0,1).  A "step" is one full solve of that system (A and b already in HBM):
low-precision LU + FP64 refinement to the FP64 backward-error target, or, for
the comparison line, FP64 LU + substitution.  TFLOP/s uses the 2n^3/3
convention.

    python bench.py [--steps K] [--warmup W] [--n N] [--mode bf16x3|fp16|bf16|fp32]

Prints one JSON line.  One GPU only: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_MFMA_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "fp16": 2500.0}
PEAK_HBM_GBS = 8000.0
MODES = {"bf16": 1, "fp16": 2, "bf16x3": 3, "fp32": 4}
MODE_PEAK = {"bf16": "bf16", "fp16": "fp16", "bf16x3": "bf16", "fp32": "f32"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e4)
    ap.add_argument("--mode", default=os.environ.get("SCIRS_BENCH_MODE", "bf16x3"), choices=sorted(MODES))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 LU comparison line")
    ap.add_argument("--cpu-n", type=int, default=2048, help="size of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def make_system(n, kappa, seed, torch):
    """A = U diag(sigma) V^T (exact kappa_2), b = A x_true, all FP64 on the GPU."""
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(seed)
    u, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    v, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    sigma = torch.logspace(0.0, float(torch.log10(torch.tensor(kappa))), n, dtype=torch.float64, device=dev)
    a = ((u * sigma) @ v.T).contiguous()
    del u, v
    x_true = torch.randn(n, 1, dtype=torch.float64, device=dev, generator=g)
    b = (a @ x_true).contiguous()
    torch.cuda.synchronize()
    return a, b, x_true


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [fn() for _ in range(steps)]
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, infos


def cpu_baseline(n_sample, seed):
    """The oracle (plain-C restatement of the reference solve, one thread)."""
    import numpy as np
    from oracle import oracle
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n_sample, n_sample))
    b = rng.standard_normal(n_sample)
    oracle.solve(a[:64, :64], b[:64])   # load / build
    t0 = time.perf_counter()
    oracle.solve(a, b)
    t = time.perf_counter() - t0
    flops = 2.0 * n_sample ** 3 / 3.0
    return {
        "value": flops / t / 1e12,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle solve (lapack.rs:72-147 + solve.rs:54-111 restated in C, 1 thread) of one "
                  f"{n_sample}x{n_sample} N(0,1) system, 1 rhs: {t:.2f} s",
        "seconds": t,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the solve engine does not shard "
              "across GPUs. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from scirs_amd import linalg as L

    n = args.n
    flops = 2.0 * n ** 3 / 3.0
    a, b, x_true = make_system(n, args.kappa, args.seed, torch)
    x = torch.empty_like(b)
    mode = MODES[args.mode]
    tol = n * 2.0 ** -53

    # ---- mixed-precision IR LU (the headline)
    def mixed():
        return L.gesv_ir_dev(a, b, x, mode=mode, tol=tol)

    for _ in range(args.warmup):
        mixed()
    torch.cuda.synchronize()
    L.prof_begin()
    t_mixed, infos = time_steps(mixed, args.steps, 0, torch)
    prof = L.prof_end()
    bwd = L.backward_error_dev(a, x, b)
    ferr = float((x - x_true).abs().max() / x_true.abs().max())
    last = infos[-1]

    # ---- plain FP64 LU at the same n
    fp64 = None
    if not args.no_fp64:
        lu = torch.empty_like(a)
        perm = torch.empty(n, dtype=torch.int64, device="cuda")
        x64 = torch.empty_like(b)

        def plain():
            return L.gesv_dev(a, b, x64, lu, perm)

        L.prof_begin()
        t64, infos64 = time_steps(plain, max(1, min(args.steps, 3)), 1, torch)
        prof64 = L.prof_end()
        bwd64 = L.backward_error_dev(a, x64, b)
        upd = prof64["update"]
        fp64 = {
            "tflops": flops / t64 / 1e12,
            "ms_per_solve": t64 * 1e3,
            "backward_error": bwd64,
            "forward_error": float((x64 - x_true).abs().max() / x_true.abs().max()),
            "factor_ms": infos64[-1].time_factor_ms,
            "update_tflops": (upd["flops"] / (upd["total_ms"] * 1e-3) / 1e12) if upd["total_ms"] else None,
            "kernel_ms": {k: round(v["total_ms"] / max(1, len(infos64)), 3) for k, v in prof64.items()},
        }
        del lu, perm, x64

    # ---- roofline of the dominant kernel class of the mixed solve
    per_step = {k: v["total_ms"] / args.steps for k, v in prof.items()}
    dom = max(per_step, key=per_step.get)
    d = prof[dom]
    avg_ms = d["total_ms"] / max(1, d["launches"])
    if dom == "update":
        achieved = d["flops"] / d["launches"] / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "trailing update (" + args.mode + " MFMA)", "achieved": achieved,
                "peak": PEAK_MFMA_TFLOPS[MODE_PEAK[args.mode]], "unit": "TFLOP/s"}
    else:
        achieved = d["bytes"] / d["launches"] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    roof["avg_launch_ms"] = avg_ms
    roof["launches_per_step"] = d["launches"] / args.steps

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_n, args.seed)

    value = flops / t_mixed / 1e12
    line = {
        "metric": "Solve TFLOP/s at FP64 backward error: mixed-prec IR LU vs plain FP64 LU",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"mixed-precision IR LU solve, n={n}, 1 rhs, FP64 input, kappa_2={args.kappa:g} "
                        f"(A = U diag(logspace) V^T), FP32 panel, {args.mode} trailing update, FP64 residual",
            "n": n, "nrhs": 1, "kappa": args.kappa, "mode": args.mode, "seed": args.seed,
        },
        "backward_error": bwd,
        "backward_error_target": tol,
        "forward_error": ferr,
        "ir_iters": last.ir_iters,
        "gmres_iters": last.gmres_iters,
        "path": last.path,
        "factor_ms": last.time_factor_ms,
        "refine_ms": last.time_solve_ms,
        "kernel_ms_per_step": {k: round(v, 3) for k, v in per_step.items()},
        "fp64_lu": fp64,
        "speedup_vs_fp64_lu": (value / fp64["tflops"]) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if bwd > tol:
        line["warning"] = f"backward error {bwd:.3e} above target {tol:.3e}: value is not at FP64 accuracy"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
