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
# algorithmic bytes of a finest-level Chebyshev step launch in FP32, mean of
# the three launches of a degree-2 V-cycle: per node x, b, dinv read and
# x_out written (4 x 3 floats) plus the clamp-mask byte; the pre-smoothing
# step (after the first step from zero, d = x) neither reads nor writes d,
# the first post-smoothing step writes it, the last one reads it: 49, 61
# and 61 bytes.  Per element the SIMP scale (float) and the 8-int connectivity.
CHEB_BYTES_PER_NODE = (49 + 61 + 61) / 3.0
CHEB_BYTES_PER_ELEM = 4 + 8 * 4
FLOPS_PER_ELEM = 1152.0      # 2 * 24 * 24
TIME_CHEB = 1 << 2           # launch-kind mask: Chebyshev update
# HBM traffic of the finest-level Chebyshev launch on the default workload
# (level 7), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (KiB per
# dispatch, mean over the launches at that level), keyed by (FP32 form, owner
# tile nodes).  The counters are scaled by what they read for a 1 GiB
# device-to-device copy in the same runs (bench.py --calibration-copy 1024):
# FETCH_SIZE 524,318.6 KiB and WRITE_SIZE 1,048,576 KiB for 1,048,576 KiB moved.
COPY_FETCH_FACTOR = 1048576.0 / 524307.1
COPY_WRITE_FACTOR = 1048576.0 / 1048576.0
MEASURED_TRAFFIC_KIB = {("valu", 1024): (173259.2, 66824.0)}
# the kernel source the counters were measured on (sha256 of
# tmr_amd/csrc/kernels.hip, first 16 hex digits): after any edit of it the
# stored traffic no longer describes the kernel and is not reported
MEASURED_TRAFFIC_SRC = "101ece533a1889d2"


def kernel_source_hash():
    import hashlib
    with open(os.path.join(ROOT, "tmr_amd", "csrc", "kernels.hip"), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(level=5):
    """The FP64 oracle's multigrid (oracle/mg_oracle.c) on the same box
    cantilever at a smaller level (bounded CPU sample): Galerkin coarse
    operators, flexible PCG with a Chebyshev degree-2 V-cycle (the GPU's FP64
    leg), and GMRES(50) with a symmetric-SOR V-cycle (TMRTopoProblem's
    solver), to 1e-8, on the host's cores (OpenMP, at most 16 threads: the
    GPU host's share) and on one core."""
    from oracle import pyoracle as po
    from tmr_amd import workloads
    forests, rho, L = workloads.box(level=level, nlevels=level - 1, seed=0)
    f0 = forests[0]
    orc = po.from_forest(f0, [("fixed", 7)])
    t0 = time.perf_counter()
    orc.assemble(rho ** 3.0 + 1e-6)
    t_asm = time.perf_counter() - t0
    f = workloads.face_load(f0, [(int(o["tag"]), int(o["info"])) for o in f0.getOctsWithName("load")],
                            [0.0, 0.0, -1.0])
    ndof = 3 * orc.nnodes
    nthreads = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for th in (nthreads, 1):
        po.set_threads(th)
        t0 = time.perf_counter()
        mg = po.Multigrid(orc, forests, [("fixed", 7)])
        mg.setup("chebyshev", degree=2, lower=1.0 / 6.0, upper=1.1)
        t_setup = time.perf_counter() - t0
        # repeated solves until about 3 s have passed (one solve is well
        # under a second at this size); the mean time per solve
        reps, t_all = 0, 0.0
        while t_all < 3.0 or reps < 3:
            t0 = time.perf_counter()
            mg.pcg(f, rtol=1e-8)
            t_all += time.perf_counter() - t0
            reps += 1
        dt = t_all / reps
        out[th] = dict(pcg_s=dt, pcg_iters=mg.iters, setup_s=t_setup, reps=reps)
        if th == nthreads:
            mg.setup("sor", omega=1.0)
            t0 = time.perf_counter()
            mg.gmres(f, m=50, rtol=1e-8)
            out[th].update(gmres_s=time.perf_counter() - t0, gmres_iters=mg.iters)
    po.set_threads(1)
    a = out[nthreads]
    return {"value": 1.0 / a["pcg_s"], "unit": "solves/s", "cores": nthreads, "kind": "port",
            "value_1core": 1.0 / out[1]["pcg_s"],
            "mdof_iter_per_s": ndof * a["pcg_iters"] / a["pcg_s"] / 1e6,
            "pcg_iters": a["pcg_iters"],
            "gmres_sor_solves_per_s": 1.0 / a["gmres_s"], "gmres_sor_iters": a["gmres_iters"],
            "sample": "oracle FP64 Galerkin MG on the same box cantilever at level %d (%d hex, %d "
                      "dofs, %d levels, coarsest 8x4x4 by dense Cholesky): flexible PCG + "
                      "Chebyshev-2 V-cycle %d iterations to 1e-8 in %.3f s on %d threads (mean "
                      "of %d solves; %.3f s on 1, mean of %d); GMRES(50) + symmetric-SOR V-cycle "
                      "%d iterations in %.2f s; assembly %.1f s and MG setup %.2f s not timed"
                      % (level, orc.nelem, ndof, len(forests), a["pcg_iters"], a["pcg_s"],
                         nthreads, a["reps"], out[1]["pcg_s"], out[1]["reps"], a["gmres_iters"],
                         a["gmres_s"], t_asm, a["setup_s"])}


def oracle_error(forest, rho, f, u, u64=None):
    """The oracle's element-by-element check of a GPU solution: true
    residual |f - K u| / |f| and the energy identity f^T u = u^T K u."""
    from oracle import pyoracle as po
    orc = po.from_forest(forest, [("fixed", 7)])
    s = rho ** 3.0 + 1e-6
    res = orc.true_residual(s, f, u)
    fr = np.where(orc.free_mask(), f, 0.0)
    c = float(fr @ u)
    out = {"true_residual": res, "energy_identity": abs(c - float(s @ orc.energy(u))) / abs(c)}
    if u64 is not None:
        out["u_rel_diff_mixed_vs_fp64"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
    return out


def bench_multirhs(args):
    """16 load cases (seeded traction patches) on the 2,160,152-hex adaptive
    forest: one batched solve of all cases against the 16 solved one by one."""
    import torch
    from tmr_amd import solver, workloads
    nrhs = args.nrhs
    t0 = time.perf_counter()
    forests, rho, L = workloads.c5a_adaptive(seed=1, nlevels=5, device=0)
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    prob.set_stream(torch.cuda.current_stream().cuda_stream)
    prob.set_design(rho)
    F, params = workloads.patch_loads(forests[0], nrhs, seed=5)
    log("multi-load-case problem built in %.1f s: %d hex, %d dofs"
        % (time.perf_counter() - t0, prob.nelem, prob.ndof))
    if not args.no_autotune:
        prob.autotune(args.tiles, reps=3)
    Fd = torch.from_numpy(F).cuda()
    Ud = torch.zeros_like(Fd)
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=400, no_graph=int(args.no_graph))
    prob.setup(opts)
    prob.solve_multi(Fd, Ud, opts=opts)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        prob.solve_multi(Fd, Ud, opts=opts)
    torch.cuda.synchronize()
    dt_b = (time.perf_counter() - ts) / args.steps
    iters_b = [i.iters for i in prob.infos]
    u1 = torch.zeros_like(Fd[0])
    prob.solve(Fd[0], u1, opts=opts)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    iters_s = []
    for k in range(nrhs):
        prob.solve(Fd[k], u1, opts=opts)
        iters_s.append(prob.info.iters)
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - ts
    ndof = prob.ndof
    out = {"metric": "load-case solves/s, %d load cases batched vs one by one" % nrhs,
           "value": nrhs / dt_b, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_step": 1e3 * dt_b, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 PCG + f32 multigrid", "data": "synthetic",
           "config": {"workload": "adaptive cantilever (2x2x2 blocks on a 2x1x1 box, levels 3-7, "
                                  "seed 1), %d seeded traction patches on its free outer faces, "
                                  "5-level MG, PCG to 1e-8" % nrhs,
                      "elements": prob.nelem, "dofs": ndof},
           "sequential_solves_per_s": nrhs / dt_s, "speedup_batched": dt_s / dt_b,
           "iters_batched": iters_b, "iters_sequential": iters_s,
           "mdof_iter_per_s": ndof * sum(iters_b) / dt_b / 1e6}
    if not args.no_oracle:
        U = Ud.cpu().numpy()
        errs = [oracle_error(forests[0], rho, F[k], U[k])["true_residual"] for k in (0, nrhs - 1)]
        out["err_vs_oracle"] = {"true_residual_cases_0_and_last": errs}
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """96 independent cantilevers of edge 8..48 elements solved side by side
    in one forest, against the same problems one by one (a sample)."""
    import torch
    from tmr_amd import solver, workloads
    t0 = time.perf_counter()
    forests, rho, edges = workloads.cube_batch(args.count, nlevels=4, device=0)
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    log("batch of %d problems built in %.1f s: %d elements, %d dofs"
        % (prob.num_components(), time.perf_counter() - t0, prob.nelem, prob.ndof))
    prob.set_stream(torch.cuda.current_stream().cuda_stream)
    prob.set_design(rho)
    if not args.no_autotune:
        prob.autotune(args.tiles, reps=3)
    f = torch.from_numpy(prob.traction_load("load", [0.0, 0.0, -1.0])).cuda()
    u = torch.zeros_like(f)
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=500, no_graph=int(args.no_graph))
    prob.setup(opts)
    prob.solve(f, u, opts=opts)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        prob.solve(f, u, opts=opts)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - ts) / args.steps
    nc = prob.num_components()
    its = [prob.component_info(k)[0] for k in range(nc)]
    # the same problems one at a time: one of each size, scaled by how often
    # the size was drawn
    seq = 0.0
    sizes = sorted(set(int(e) for e in edges))
    for e in sizes:
        k = [int(x) for x in edges].index(e)
        fk, rk = workloads.uneven_cube_batch([e], nlevels=4, seed=k, device=0)
        pk = solver.Problem(fk, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
        pk.set_stream(torch.cuda.current_stream().cuda_stream)
        pk.set_design(rk)
        gk = torch.from_numpy(pk.traction_load("load", [0.0, 0.0, -1.0])).cuda()
        uk = torch.zeros_like(gk)
        pk.setup(opts)
        pk.solve(gk, uk, opts=opts)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        pk.solve(gk, uk, opts=opts)
        torch.cuda.synchronize()
        seq += (time.perf_counter() - ts) * sum(1 for x in edges if int(x) == e)
    seq_all = seq
    out = {"metric": "independent problems solved/s, %d uneven cantilevers in one launch per "
                     "stage vs one by one" % nc,
           "value": nc / dt, "unit": "problems/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 PCG + f32 multigrid", "data": "synthetic",
           "config": {"workload": "%d cubes of seeded edges 8..48 hex8 (every integer equally "
                                  "likely), 4-level MG per cube (n, ceil(n/2), ...; coarse levels "
                                  "not nested where n has no coarsen() level), PCG to 1e-8 per "
                                  "problem" % nc,
                      "edge_counts": {str(e): sum(1 for x in edges if int(x) == e)
                                      for e in sorted(set(int(x) for x in edges))},
                      "elements": prob.nelem, "dofs": prob.ndof},
           "sequential_problems_per_s": nc / seq_all,
           "speedup_batched": seq_all / dt,
           "sequential_sample": "each drawn size solved alone once, times its count",
           "iters_min_max": [min(its), max(its)]}
    if not args.no_oracle:
        # per problem: the oracle's element-by-element true residual over that
        # cube's dofs (cube k holds (n_k + 1)^3 consecutive nodes)
        from oracle import pyoracle as po
        t0 = time.perf_counter()
        orc = po.from_forest(forests[0], [("fixed", 7)])
        uh, fh = u.cpu().numpy(), f.cpu().numpy()
        fr = np.where(orc.free_mask(), fh, 0.0)
        r = fr - orc.apply_free(rho ** 3.0 + 1e-6, uh)
        bounds = np.concatenate([[0], np.cumsum([3 * (int(e) + 1) ** 3 for e in edges])])
        assert bounds[-1] == len(fh)
        res = [float(np.linalg.norm(r[a:b]) / np.linalg.norm(fr[a:b]))
               for a, b in zip(bounds[:-1], bounds[1:])]
        out["err_vs_oracle"] = {"per_problem_true_residual_max": max(res),
                                "per_problem_true_residual_median": float(np.median(res)),
                                "what": "|f - K u| / |f| per cube, K the oracle's FP64 "
                                        "element-by-element operator"}
        log("oracle check %.1f s" % (time.perf_counter() - t0))
    print(json.dumps(out), flush=True)


def bench_adaptive(args):
    """The adaptive cantilever (1,227,458 hex, levels 3-7, hanging nodes,
    32-Gaussian density, 5-level MG): FP64 end to end, rediscretised and
    Galerkin coarse levels, and the FP32 V-cycle, to 1e-8."""
    import torch
    from tmr_amd import solver, workloads
    t0 = time.perf_counter()
    forests, rho, L = workloads.c2_adaptive(seed=0, nlevels=5, device=0)
    log("adaptive forest hierarchy %s built in %.1f s" % ([len(x.getOctants()) for x in forests],
                                                           time.perf_counter() - t0))
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    prob.set_stream(torch.cuda.current_stream().cuda_stream)
    prob.set_design(rho)
    f_host = prob.traction_load("load", [0.0, 0.0, -1.0])
    f = torch.from_numpy(f_host).cuda()
    u = torch.zeros_like(f)
    if not args.no_autotune:
        prob.autotune(args.tiles, reps=3)
    legs = {}
    plan = [("fp64", 0, False, False), ("fp64_galerkin", 0, True, False),
            ("mixed", 1, False, False), ("mixed_galerkin", 1, True, False)]
    if args.sor:  # TMR's default smoother, symmetric SOR on the assembled levels
        plan += [("fp64_sor", 0, False, True), ("mixed_sor", 1, False, True),
                 ("mixed_galerkin_sor", 1, True, True)]
    for name, prec, gal, sor in plan:
        prob.set_multigrid(use_galerkin=gal, use_chebyshev_smoother=not sor)
        opts = solver.default_opts(rtol=1e-8, precision=prec, max_iters=400,
                                   no_graph=int(args.no_graph))
        torch.cuda.synchronize()
        ts = time.perf_counter()
        prob.setup(opts)
        torch.cuda.synchronize()
        setup_ms = 1e3 * (time.perf_counter() - ts)
        for _ in range(args.warmup):
            prob.solve(f, u, opts=opts)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        its = []
        for _ in range(args.steps):
            prob.solve(f, u, opts=opts)
            its.append(prob.info.iters)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - ts) / args.steps
        legs[name] = {"solves_per_s": 1.0 / dt, "ms_per_solve": 1e3 * dt, "pcg_iters": its[-1],
                      "setup_ms": setup_ms, "true_residual": prob.true_residual(),
                      "compliance": prob.compliance(),
                      "mdof_iter_per_s": prob.ndof * its[-1] / dt / 1e6}
        log("%s: %s" % (name, legs[name]))
    u_host = u.cpu().numpy()
    best = max(legs, key=lambda k: legs[k]["solves_per_s"])
    out = {"metric": "adaptive-forest solves/s to 1e-8 (FP64 end to end and mixed), 1.2M-hex "
                     "SIMP", "value": legs["fp64"]["solves_per_s"], "unit": "solves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": legs["fp64"]["ms_per_solve"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "adaptive cantilever, 2x2x2 blocks on a 2x1x1 box, levels 3-7 "
                                  "(seed 0), 32-Gaussian density in [1e-3, 1], 5-level MG",
                      "elements": prob.nelem, "dofs": prob.ndof,
                      "hex_per_level": [len(x.getOctants()) for x in forests]},
           "legs": legs, "fastest_leg": best}
    if not args.no_oracle:
        out["err_vs_oracle"] = oracle_error(forests[0], rho, f_host, u_host)
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Backward of the adaptive cantilever (C2 mesh, FP32 V-cycle): the
    compliance gradient and the p-norm (P = 8) stress gradient with its
    adjoint solve; time and extra device memory of each backward."""
    import torch
    from tmr_amd import solver, workloads
    forests, rho, L = workloads.c2_adaptive(seed=0, nlevels=5, device=0)
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    prob.set_stream(torch.cuda.current_stream().cuda_stream)
    prob.set_design(rho)
    f = torch.from_numpy(prob.traction_load("load", [0.0, 0.0, -1.0])).cuda()
    u = torch.zeros_like(f)
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=400, no_graph=int(args.no_graph))
    prob.setup(opts)
    g = torch.empty(prob.nelem, dtype=torch.float64, device="cuda")
    times = {"forward_solve": [], "compliance_backward": [], "stress_value": [],
             "stress_backward": []}
    mem = {}
    for step in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        marks = [time.perf_counter()]
        prob.solve(f, u, opts=opts)
        torch.cuda.synchronize()
        marks.append(time.perf_counter())
        prob.compliance_grad(g)
        torch.cuda.synchronize()
        marks.append(time.perf_counter())
        prob.stress_pnorm(8.0)
        marks.append(time.perf_counter())
        prob.stress_pnorm_grad(8.0, g, opts=opts)
        torch.cuda.synchronize()
        marks.append(time.perf_counter())
        mem["extra_device_mib_first_step" if step == 0 else "extra_device_mib"] = \
            (free0 - torch.cuda.mem_get_info()[0]) / 2 ** 20
        if step >= args.warmup:
            for k, a, b in zip(times, marks[:-1], marks[1:]):
                times[k].append(1e3 * (b - a))
    ms = {k: float(np.mean(v)) for k, v in times.items()}
    step_ms = sum(ms.values())
    out = {"metric": "differentiation step (solve + compliance and p-norm gradients) per second, "
                     "1.2M-hex adaptive", "value": 1e3 / step_ms, "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 PCG + f32 multigrid", "data": "synthetic",
           "config": {"workload": "adaptive cantilever (C2 mesh), FP32 V-cycle, compliance and "
                                  "p-norm (P = 8) von Mises gradients",
                      "elements": prob.nelem, "dofs": prob.ndof},
           "ms": ms, "memory": mem, "adjoint_iters": prob.adjoint_info.iters,
           "forward_iters": prob.info.iters,
           "device_memory_total_mib": (torch.cuda.mem_get_info()[1]
                                       - torch.cuda.mem_get_info()[0]) / 2 ** 20}
    print(json.dumps(out), flush=True)


def bench_forest(args):
    """Forest setup of the two benchmark meshes with the host passes and with
    the GPU passes (balance, createNodes, createInterpolation): the 6-level
    box hierarchy (4.2M hex at the finest level) and the 5-level adaptive
    hierarchy (1.2M hex, its refinement sweeps included), each checked
    bit-exact against the host build."""
    import numpy as np

    from tmr_amd import workloads

    def build(kind, device):
        t = time.perf_counter()
        if kind == "box":
            forests, _, _ = workloads.box(level=7, nlevels=6, seed=0, device=device)
        else:
            forests, _, _ = workloads.c2_adaptive(seed=0, nlevels=5, device=device)
        t_forest = time.perf_counter() - t
        t = time.perf_counter()
        P = [forests[l].createInterpolation(forests[l + 1], device=device)
             for l in range(len(forests) - 1)]
        return forests, P, t_forest, time.perf_counter() - t

    out = {}
    for kind in ("box", "adaptive"):
        fh, Ph, th, tih = build(kind, None)
        fd, Pd, td, tid = build(kind, 0)
        same = all(np.array_equal(a.getOctants(), b.getOctants()) and
                   np.array_equal(a.getMeshConn(), b.getMeshConn()) and
                   all(np.array_equal(x, y) for x, y in zip(a.getDepNodeConn(), b.getDepNodeConn()))
                   and np.array_equal(a.getPoints(), b.getPoints()) for a, b in zip(fh, fd))
        same = same and all(np.array_equal(x, y) for a, b in zip(Ph, Pd) for x, y in zip(a, b))
        out[kind] = {"hex": len(fh[0].getOctants()), "levels": len(fh),
                     "host_s": {"forests": th, "interpolation": tih, "total": th + tih},
                     "gpu_s": {"forests": td, "interpolation": tid, "total": td + tid},
                     "speedup": (th + tih) / (td + tid), "bit_exact": bool(same)}
        log(kind, out[kind])
    res = {"metric": "forest hierarchy setups/s (balance, nodes, interpolation on the GPU), "
                     "4.2M-hex box", "value": 1.0 / out["box"]["gpu_s"]["total"],
           "unit": "setups/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": 1e3 * out["box"]["gpu_s"]["total"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32 octants, f64 weights",
           "data": "synthetic",
           "config": {"workload": "6-level box hierarchy (256x128x128 hex8 finest) and the "
                                  "5-level adaptive hierarchy (1,227,458 hex), host vs GPU "
                                  "forest passes"},
           "meshes": out}
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--nlevels", type=int, default=6)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the oracle's residual check of the solution")
    ap.add_argument("--cpu-level", type=int, default=5)
    ap.add_argument("--fp32-valu", action="store_true",
                    help="FP32 operator on the vector ALUs instead of the matrix cores")
    ap.add_argument("--fp32-form", type=int, choices=(0, 1, 2), default=None,
                    help="pin the FP32 operator form after the autotuner: 0 vector ALUs, 1 matrix "
                         "cores (bf16 split), 2 matrix cores (FP32 operands)")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue every PCG iteration instead of replaying HIP graphs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default owner-tile size")
    ap.add_argument("--workload", default="cantilever",
                    choices=["cantilever", "multirhs", "batch", "adaptive", "adjoint", "forest"],
                    help="cantilever: the 4.2M-hex headline; multirhs: 16 load cases on a "
                         "2.1M-hex cube; batch: 96 uneven independent cantilevers; adaptive: "
                         "the 1.2M-hex adaptive forest; adjoint: its gradients; forest: host vs "
                         "GPU forest setup")
    ap.add_argument("--sor", action="store_true",
                    help="adaptive workload: add legs with symmetric SOR smoothing")
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--calibration-copy", type=int, default=0, metavar="MIB",
                    help="after the solves, copy MIB MiB device to device (a known-size "
                         "reference for memory-traffic counters)")
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--tiles", default="auto",
                    help="owner-tile sizes (nodes) the autotuner tries, comma-separated "
                         "(auto: Problem.tile_candidates())")
    args = ap.parse_args()
    args.tiles = None if args.tiles == "auto" else tuple(int(v) for v in args.tiles.split(","))
    if args.gpus != 1:
        print("bench.py: the solver runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "multirhs":
        return bench_multirhs(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "adaptive":
        return bench_adaptive(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "forest":
        return bench_forest(args)

    import torch
    from tmr_amd import solver, workloads

    t0 = time.perf_counter()
    forests, rho, L = workloads.box(level=args.level, nlevels=args.nlevels, seed=0, device=0)
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
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = prob.autotune(args.tiles, reps=3)
        log("autotuned owner tiles: %d nodes, FP32 operator on the %s (ms per 3 FP32 Chebyshev "
            "steps + FP32 and FP64 residuals: %s) in %.1f s"
            % (tune[0], solver.Problem.FP32_FORMS[prob.tuning()[1]], tune[1],
               time.perf_counter() - t0))
    if args.fp32_form is not None:
        prob.set_fp32_form(args.fp32_form)

    rho_dev = torch.from_numpy(rho).cuda()

    def run(precision, steps, warmup, timing=False):
        opts = solver.default_opts(rtol=1e-8, precision=precision, max_iters=500,
                                   fp32_valu=int(args.fp32_valu), no_graph=int(args.no_graph))
        torch.cuda.synchronize()
        ts = time.perf_counter()
        prob.setup(opts)
        torch.cuda.synchronize()
        first_setup_ms = 1e3 * (time.perf_counter() - ts)
        # a design step of an optimisation loop: the next design (resident
        # on the device) installed and the preconditioner set up again,
        # without the first setup's one-time allocations
        ts = time.perf_counter()
        prob.set_design(rho_dev)
        prob.setup(opts)
        torch.cuda.synchronize()
        setup_ms = 1e3 * (time.perf_counter() - ts)
        for _ in range(warmup):
            prob.solve(f, u, opts=opts)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        iters = []
        for _ in range(steps):
            prob.solve(f, u, opts=opts)
            iters.append(prob.info.iters)
        torch.cuda.synchronize()
        dt = time.perf_counter() - ts
        res = dict(seconds=dt, iters=iters, setup_ms=setup_ms, first_setup_ms=first_setup_ms,
                   true_res=prob.true_residual(), compliance=prob.compliance(),
                   u=u.cpu().numpy().copy())
        if timing:
            # kernel durations from one more (untimed, ungraphed) solve: HIP
            # events around every finest-level Chebyshev-step launch
            prob.kernel_timing(TIME_CHEB)
            prob.solve(f, u, opts=opts)
            res["kt"] = prob.kernel_timing_result()
            prob.kernel_timing(False)
            # host buffers: f copied in and u copied out around each solve
            u_host = np.zeros_like(f_host)
            prob.solve(f_host, u_host, opts=opts)
            ts = time.perf_counter()
            for _ in range(steps):
                prob.solve(f_host, u_host, opts=opts)
            res["host_seconds"] = time.perf_counter() - ts
        return res

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
    bytes_launch = int(round(CHEB_BYTES_PER_NODE * nnodes + CHEB_BYTES_PER_ELEM * prob.nelem))
    achieved = bytes_launch / avg_s / 1e9
    form = ("valu", "mfma", "mfma32")[0 if args.fp32_valu else prob.tuning()[1]]
    traffic = None
    pmc = MEASURED_TRAFFIC_KIB.get((form, prob.tuning()[0])) if args.level == 7 else None
    src = kernel_source_hash()
    if src != MEASURED_TRAFFIC_SRC:
        pmc = None
    if pmc is not None:
        traffic = 1024.0 * (pmc[0] * COPY_FETCH_FACTOR + pmc[1] * COPY_WRITE_FACTOR)
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
                               "k0=1e-6, 6-level MG (Chebyshev deg 2 on [lambda/6, 1.1 lambda]), "
                               "FP64 flexible PCG to 1e-8" % args.level,
                   "elements": prob.nelem, "dofs": ndof, "mg_levels": prob.nlevels},
        "mdof_iter_per_s": mdof,
        "host_buffer_solves_per_s": args.steps / mixed["host_seconds"],
        "pcg_iters": mixed["iters"],
        "true_residual": mixed["true_res"],
        "compliance": mixed["compliance"],
        "setup_ms": mixed["setup_ms"],
        "first_setup_ms": mixed["first_setup_ms"],
        # a design step of an optimisation loop: the design installed and the
        # density-dependent setup (diagonals, eigenvalue estimates, coarse
        # factor) redone, plus one solve
        "designs_per_s": 1.0 / (mixed["setup_ms"] / 1e3 + mixed["seconds"] / args.steps),
        "tile_nodes": prob.tuning()[0],
        "fp32_operator": {"valu": "vector ALUs", "mfma": "matrix cores (bf16 three-way split)",
                          "mfma32": "matrix cores (FP32 operands)"}[form],
        "hip_graphs": not args.no_graph,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": {"valu": "k_tile_apply<float, ASM_CHEB>",
                                "mfma": "k_tile_apply_mfma<ASM_CHEB>",
                                "mfma32": "k_tile_apply_mfma32<ASM_CHEB>"}[form]
                               + " (finest-level Chebyshev step)",
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE measured on kernels.hip "
                                       "%s; this run's kernels.hip %s" % (MEASURED_TRAFFIC_SRC, src),
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
    if not args.no_oracle:
        # the oracle's own check of the mixed leg's solution (not timed)
        log("oracle check of the solution ...")
        t0 = time.perf_counter()
        out["err_vs_oracle"] = oracle_error(prob.forests[0], rho, f_host, mixed["u"],
                                            fp64["u"] if fp64 is not None else None)
        out["err_vs_oracle"]["what"] = ("FP32-V-cycle leg's u: true residual |f - K u| / |f| and "
                                        "|f^T u - u^T K u| / f^T u from the oracle's "
                                        "element-by-element FP64 operator; u difference to the "
                                        "FP64 leg")
        log("oracle check %.1f s: %s" % (time.perf_counter() - t0, out["err_vs_oracle"]))
    if args.calibration_copy > 0:
        n = args.calibration_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = 4 * n
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_level)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
