#!/usr/bin/env python3
"""Benchmark: solves/s and MDOF*iter/s of the multigrid-preconditioned FP64
PCG on the 2 x 1 x 1 box cantilever (256 x 128 x 128 = 4,194,304 hex8,
12,830,211 dofs, 6-level hierarchy, coarsest 8 x 4 x 4 by FP64 Cholesky),
FP32 V-cycle (value) and FP64 V-cycle (comparison), to a 1e-8 relative
residual.  One step = one solve K(rho) u = f from u = 0 with f and u resident
in HBM; the density-dependent multigrid setup is timed separately.

Prints one JSON line (see the keys at the end of main()).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic bytes of one finest-level Chebyshev step launch in FP32: per node
# x, b, d, dinv read and d, x_out written (6 x 3 floats) plus the clamp mask
# byte; per element the SIMP scale (float) and the 8-int connectivity
CHEB_BYTES_PER_NODE = 6 * 3 * 4 + 1
CHEB_BYTES_PER_ELEM = 4 + 8 * 4
FLOPS_PER_ELEM = 1152.0      # 2 * 24 * 24
TIME_CHEB = 1 << 2           # launch-kind mask: Chebyshev update


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(level=5, iters=None):
    """FP64 oracle (assembled CSR + Jacobi PCG, 1 core) on the same box
    cantilever at a smaller level; reports MDOF*iter/s."""
    from oracle import pyoracle as po
    from tmr_amd import workloads
    forests, rho, L = workloads.box(level=level, nlevels=1, seed=0)
    f0 = forests[0]
    orc = po.from_forest(f0, [("fixed", 7)])
    orc.assemble(rho ** 3.0 + 1e-6)
    conn = f0.getMeshConn()
    ext = f0.getElementExtents()
    f = np.zeros(3 * orc.nnodes)
    for o in f0.getOctsWithName("load"):
        e, face = o["tag"], o["info"]
        ax = face // 2
        area = ext[e][(ax + 1) % 3] * ext[e][(ax + 2) % 3]
        for k in range(8):
            bit = [(k & 1), (k >> 1) & 1, (k >> 2)][ax]
            if bit == face % 2:
                f[3 * conn[e][k] + 2] += -0.25 * area
    t0 = time.perf_counter()
    try:
        orc.solve(f, rtol=1e-8, maxit=iters or 20000)
        it = orc.iters
    except RuntimeError:
        it = iters
    dt = time.perf_counter() - t0
    ndof = 3 * orc.nnodes
    return {"value": ndof * it / dt / 1e6, "unit": "MDOF*iter/s", "cores": 1, "kind": "port",
            "sample": "oracle FP64 Jacobi-PCG (assembled CSR) on the same box cantilever at "
                      "level %d (%d hex, %d dofs), %d iterations to 1e-8 in %.1f s"
                      % (level, conn.shape[0], ndof, it, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--nlevels", type=int, default=6)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-level", type=int, default=5)
    ap.add_argument("--fp32-valu", action="store_true",
                    help="FP32 operator on the vector ALUs instead of the matrix cores")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from tmr_amd import solver, workloads

    t0 = time.perf_counter()
    forests, rho, L = workloads.box(level=args.level, nlevels=args.nlevels, seed=0)
    log("forest hierarchy built in %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    log("problem (CSR flatten + upload) built in %.1f s: %d elements, %d dofs"
        % (time.perf_counter() - t0, prob.nelem, prob.ndof))
    stream = torch.cuda.current_stream()
    prob.set_stream(stream.cuda_stream)
    f_host = prob.traction_load("load", [0.0, 0.0, -1.0])
    f = torch.from_numpy(f_host).cuda()
    u = torch.zeros_like(f)
    prob.set_design(rho)

    def run(precision, steps, warmup, timing=False):
        opts = solver.default_opts(rtol=1e-8, precision=precision, max_iters=500,
                                   fp32_valu=int(args.fp32_valu))
        torch.cuda.synchronize()
        ts = time.perf_counter()
        prob.setup(opts)
        torch.cuda.synchronize()
        setup_ms = 1e3 * (time.perf_counter() - ts)
        for _ in range(warmup):
            prob.solve(f, u, opts=opts)
        if timing:
            prob.kernel_timing(TIME_CHEB)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        iters = []
        for _ in range(steps):
            prob.solve(f, u, opts=opts)
            iters.append(prob.info.iters)
        torch.cuda.synchronize()
        dt = time.perf_counter() - ts
        kt = prob.kernel_timing_result() if timing else None
        if timing:
            prob.kernel_timing(False)
        return dict(seconds=dt, iters=iters, setup_ms=setup_ms, kt=kt,
                    true_res=prob.true_residual(), compliance=prob.compliance())

    mixed = run(1, args.steps, args.warmup, timing=True)
    log("mixed: %.3f s for %d solves, iters %s, true residual %.2e"
        % (mixed["seconds"], args.steps, mixed["iters"], mixed["true_res"]))
    fp64 = None
    if not args.no_fp64:
        fp64 = run(0, args.steps, args.warmup)
        log("fp64: %.3f s for %d solves, iters %s, true residual %.2e"
            % (fp64["seconds"], args.steps, fp64["iters"], fp64["true_res"]))

    ndof = prob.ndof
    solves_s = args.steps / mixed["seconds"]
    mdof = ndof * sum(mixed["iters"]) / mixed["seconds"] / 1e6
    tot_ms, cnt, _ = mixed["kt"]
    avg_s = tot_ms / max(cnt, 1) / 1e3
    nnodes = ndof // 3
    bytes_launch = CHEB_BYTES_PER_NODE * nnodes + CHEB_BYTES_PER_ELEM * prob.nelem
    achieved = bytes_launch / avg_s / 1e9
    out = {
        "metric": "solves/s + MDOF*iter/s: mixed-prec vs FP64 MG-PCG to 1e-8, 4.2M-hex SIMP",
        "value": solves_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 PCG + f32 multigrid",
        "data": "synthetic (seeded smoothed densities, unit traction)",
        "config": {"workload": "box cantilever 256x128x128 hex8 (2 blocks, level %d), SIMP q=3 "
                               "k0=1e-6, 6-level MG (Chebyshev deg 2), FP64 flexible PCG to 1e-8"
                               % args.level,
                   "elements": prob.nelem, "dofs": ndof, "mg_levels": prob.nlevels},
        "mdof_iter_per_s": mdof,
        "pcg_iters": mixed["iters"],
        "true_residual": mixed["true_res"],
        "compliance": mixed["compliance"],
        "setup_ms": mixed["setup_ms"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": ("k_tile_apply<float, ASM_CHEB>" if args.fp32_valu else
                                "k_tile_apply_mfma<ASM_CHEB>") + " (finest-level Chebyshev step)",
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "avg_launch_us": avg_s * 1e6, "launches": cnt,
                     "mfma_or_valu_tflops": FLOPS_PER_ELEM * prob.nelem / avg_s / 1e12,
                     "fp32_peak_tflops": FP32_PEAK_TFLOPS},
    }
    if fp64 is not None:
        out["fp64"] = {"solves_per_s": args.steps / fp64["seconds"],
                       "mdof_iter_per_s": ndof * sum(fp64["iters"]) / fp64["seconds"] / 1e6,
                       "pcg_iters": fp64["iters"], "true_residual": fp64["true_res"],
                       "compliance": fp64["compliance"], "setup_ms": fp64["setup_ms"]}
        out["speedup_mixed_vs_fp64"] = solves_s / out["fp64"]["solves_per_s"]
        out["compliance_rel_diff_mixed_vs_fp64"] = abs(mixed["compliance"] - fp64["compliance"]) / abs(fp64["compliance"])
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_level)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
