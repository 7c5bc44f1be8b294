"""Every benchmark configuration at its full size on the GPU, checked by the
oracle through size-independent properties (the oracle cannot assemble or
solve these sizes in seconds; the same configurations at oracle-solvable
sizes are compared solution for solution in test_gpu_solver.py,
test_gpu_adjoint.py, test_gpu_batch.py and test_gpu_galerkin.py):

* the true residual |f - K u| / |f| computed by the oracle's own
  element-by-element operator (oracle/fe_oracle.c orc_apply_free), not the
  GPU's;
* the energy identity f^T u = u^T K u with u^T K u summed from the oracle's
  element energies;
* FP32-V-cycle against FP64-V-cycle solutions (compliance and u);
* gradients against the oracle's formula evaluated on the GPU's solution,
  and central differences at 64 seeded elements;
* per-problem oracle residuals and, for the smaller problems of the batch,
  the oracle's own solves.

C1 (the 16^3 cube, rho = 0.5, 3 levels, rtol 1e-10) is test_gpu_solver.py's
cube: compliance and u against the oracle's solve.
"""
import time

import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6


def _oracle_checks(forest, rho, f, u, what):
    """(oracle true residual, energy-identity defect) of a solution."""
    t = time.time()
    orc = po.from_forest(forest, BCS)
    s = rho ** Q + K0
    res = orc.true_residual(s, f, u)
    e = orc.energy(u)
    fr = np.where(orc.free_mask(), f, 0.0)
    c = float(fr @ u)
    ident = abs(c - float(s @ e)) / abs(c)
    print("%s: oracle true residual %.3e, |f^T u - u^T K u| / f^T u = %.2e (%.1f s)"
          % (what, res, ident, time.time() - t), flush=True)
    return res, ident, orc


# ---------------------------------------------------------------------------
# C2: adaptive forest, about 1M hex (1,227,458), 5-level MG, FP64 end to end
@pytest.fixture(scope="module")
def c2():
    t = time.time()
    forests, rho, L = workloads.c2_adaptive(seed=0, nlevels=5)
    print("C2 forests: %s hex per level (%.1f s)" % ([len(f.getOctants()) for f in forests],
                                                     time.time() - t), flush=True)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    return forests, rho, prob, f


@pytest.mark.parametrize("galerkin", [False, True])
def test_c2_fp64_solve(c2, galerkin):
    forests, rho, prob, f = c2
    prob.set_multigrid(use_galerkin=galerkin)
    opts = solver.default_opts(rtol=1e-8, precision=0, max_iters=400)
    t = time.time()
    prob.setup(opts)
    t_setup = time.time() - t
    u = prob.solve(f, opts=opts)
    t = time.time()
    prob.solve(f, u, opts=opts)
    t_solve = time.time() - t
    print("C2 FP64 (%s coarse levels): %d iterations, setup %.1f ms, solve %.1f ms, GPU true "
          "residual %.2e" % ("Galerkin" if galerkin else "rediscretised", prob.info.iters,
                             1e3 * t_setup, 1e3 * t_solve, prob.true_residual()), flush=True)
    res, ident, _ = _oracle_checks(forests[0], rho, f, u, "C2")
    assert prob.info.converged
    assert res < 1e-8
    assert ident < 1e-6
    # regression bound on the iteration count (measured: 24 rediscretised, 20
    # Galerkin; the same mesh at uniform density takes 17, a uniform mesh of
    # smoothed densities 10)
    assert prob.info.iters <= 30
    prob.set_multigrid(use_galerkin=False)


# ---------------------------------------------------------------------------
# C4: gradients on the C2 mesh, mixed-precision solve
def test_c4_compliance_gradient(c2):
    forests, rho, prob, f = c2
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=400)
    u = prob.solve(f, opts=opts)
    import torch
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t = time.time()
    g = prob.compliance_grad()
    t_bwd = time.time() - t
    res, ident, orc = _oracle_checks(forests[0], rho, f, u, "C4 mixed solve")
    assert res < 1e-8
    # the oracle's gradient formula (-p'(rho_e) u_e^T K_e u_e) on the same u
    g_ref = -np.array([po.dpenalty(r, Q) for r in rho]) * orc.energy(u)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("C4 compliance gradient vs oracle formula: rel err %.2e; backward %.1f ms, device "
          "memory in use %.0f MiB" % (err, 1e3 * t_bwd,
                                      (torch.cuda.mem_get_info()[1] - free0) / 2 ** 20), flush=True)
    assert err < 1e-10
    # and against the gradient of the FP64 solve (solve error only)
    prob.solve(f, opts=solver.default_opts(rtol=1e-10, precision=0, max_iters=400))
    g64 = prob.compliance_grad()
    e2 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    print("C4 gradient of the mixed solve vs FP64 solve: rel diff %.2e" % e2, flush=True)
    assert e2 < 1e-6


def _fd_elements(n, count=64, seed=11):
    return np.sort(np.random.default_rng(seed).choice(n, size=count, replace=False))


def test_c4_central_differences(c2):
    """Central differences of compliance and of the p-norm (P = 8) at 64
    seeded elements, step 1e-6, FP64 solves to rtol 1e-12.  A one-element
    change of most elements moves these functionals by less than the
    accuracy of the solves (about 1e-13 of their value; the p-norm is
    dominated by a few highly stressed elements), so every element is held
    to |fd - g| <= 1e-3 |g| + 100 x that noise floor: elements above the
    floor agree to the relative bound (measured: compliance 2e-5, p-norm
    1e-6), the rest to the absolute one.  The same 64 elements moved
    together along a random direction (step 1e-4) give a change far above
    the floor, held to 1e-4."""
    forests, rho, prob, f = c2
    opts = solver.default_opts(rtol=1e-12, precision=0, max_iters=600)
    prob.set_design(rho)
    prob.solve(f, opts=opts)
    c0 = prob.compliance()
    gc = prob.compliance_grad()
    gs = prob.stress_pnorm_grad(8.0, opts=opts)
    s0 = prob.stress_pnorm(8.0)
    els = _fd_elements(prob.nelem)
    h = 1e-6
    fdc, fds = [], []
    t = time.time()
    for e in els:
        vals = []
        for sgn in (1.0, -1.0):
            r = rho.copy()
            r[e] += sgn * h
            prob.set_design(r)
            prob.solve(f, opts=opts)
            vals.append((prob.compliance(), prob.stress_pnorm(8.0)))
        fdc.append((vals[0][0] - vals[1][0]) / (2 * h))
        fds.append((vals[0][1] - vals[1][1]) / (2 * h))
    prob.set_design(rho)
    fdc, fds = np.array(fdc), np.array(fds)
    ec = np.abs(fdc - gc[els]) / np.abs(gc[els])
    es = np.abs(fds - gs[els]) / np.maximum(np.abs(gs[els]), 1e-300)
    # difference noise: relative accuracy ~1e-13 of the functional over 2h
    floor_c = 1e-13 * abs(c0) / (2 * h)
    floor_s = 1e-13 * abs(s0) / (2 * h)
    big_c = np.abs(gc[els]) > 100 * floor_c
    big_s = np.abs(gs[els]) > 100 * floor_s
    print("C4 FD (h = 1e-6, 64 elements, %.0f s): compliance rel err median %.2e max %.2e "
          "(%d above the noise floor, max there %.2e); p-norm rel err median %.2e max %.2e "
          "(%d above the floor, max there %.2e)"
          % (time.time() - t, np.median(ec), ec.max(), big_c.sum(),
             ec[big_c].max() if big_c.any() else 0.0, np.median(es), es.max(), big_s.sum(),
             es[big_s].max() if big_s.any() else 0.0), flush=True)
    assert np.all(np.abs(fdc - gc[els]) <= 1e-3 * np.abs(gc[els]) + 100 * floor_c)
    assert np.all(np.abs(fds - gs[els]) <= 1e-3 * np.abs(gs[els]) + 100 * floor_s)
    assert big_c.sum() >= 1 and big_s.sum() >= 1
    # all 64 at once along a random direction: a change far above the floor
    d = np.zeros(prob.nelem)
    d[els] = np.random.default_rng(3).standard_normal(len(els))
    hd = 1e-4
    vals = []
    for sgn in (1.0, -1.0):
        prob.set_design(rho + sgn * hd * d)
        prob.solve(f, opts=opts)
        vals.append((prob.compliance(), prob.stress_pnorm(8.0)))
    prob.set_design(rho)
    dc = (vals[0][0] - vals[1][0]) / (2 * hd)
    ds = (vals[0][1] - vals[1][1]) / (2 * hd)
    print("C4 directional FD over the 64 elements: compliance %.10e vs %.10e, p-norm %.10e vs "
          "%.10e" % (dc, gc @ d, ds, gs @ d), flush=True)
    assert abs(dc - gc @ d) / abs(gc @ d) < 1e-4
    assert abs(ds - gs @ d) / abs(gs @ d) < 1e-4


def test_c4_stress_value_and_adjoint(c2):
    import torch
    forests, rho, prob, f = c2
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=400)
    prob.set_design(rho)
    u = prob.solve(f, opts=opts)
    val = prob.stress_pnorm(8.0)
    orc = po.from_forest(forests[0], BCS)
    ref = po.stress_pnorm(orc, rho, u, P=8.0)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t = time.time()
    g = prob.stress_pnorm_grad(8.0, opts=opts)
    t_bwd = time.time() - t
    used = (free0 - torch.cuda.mem_get_info()[0]) / 2 ** 20
    print("C4 p-norm %.12g oracle %.12g; backward (adjoint solve %d iterations + kernels) "
          "%.1f ms, extra device memory %.0f MiB"
          % (val, ref, prob.adjoint_info.iters, 1e3 * t_bwd, used), flush=True)
    assert abs(val - ref) / abs(ref) < 1e-10
    assert prob.adjoint_info.converged
    assert np.all(np.isfinite(g))


# ---------------------------------------------------------------------------
# C3: 256 x 128 x 128 box (4,194,304 hex), FP32 V-cycle inside FP64 PCG
def test_c3_mixed_vs_fp64():
    t = time.time()
    forests, rho, L = workloads.box(level=7, nlevels=6, seed=0)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    print("C3 built: %d hex, %d dofs (%.1f s)" % (prob.nelem, prob.ndof, time.time() - t),
          flush=True)
    out = {}
    for prec in (1, 0):
        opts = solver.default_opts(rtol=1e-8, precision=prec, max_iters=200)
        u = prob.solve(f, opts=opts)
        out[prec] = (u, prob.compliance(), prob.info.iters, prob.true_residual())
        print("C3 %s: %d iterations, GPU true residual %.3e, compliance %.15g"
              % ("mixed" if prec else "FP64", out[prec][2], out[prec][3], out[prec][1]),
              flush=True)
    (um, cm, im, rm), (ud, cd, i64, r64) = out[1], out[0]
    dc = abs(cm - cd) / abs(cd)
    du = np.linalg.norm(um - ud) / np.linalg.norm(ud)
    print("C3 mixed vs FP64: compliance rel diff %.2e, u rel diff %.2e" % (dc, du), flush=True)
    for u, what in ((um, "C3 mixed"), (ud, "C3 FP64")):
        res, ident, _ = _oracle_checks(forests[0], rho, f, u, what)
        assert res < 1e-8
        assert ident < 1e-7
    assert dc < 1e-6 and du < 1e-6
    assert im <= 12 and i64 <= 12
    # the same problem on 512-node owner tiles (set_tiling): same answer
    prob.set_tiling(512)
    nt, tn = prob.tile_info(0)
    print("C3 on %d tiles of %d nodes" % (nt, tn), flush=True)
    assert tn == 512 and nt == 8354
    us = prob.solve(f, opts=solver.default_opts(rtol=1e-8, precision=1, max_iters=200))
    its, trs = prob.info.iters, prob.true_residual()
    res, ident, _ = _oracle_checks(forests[0], rho, f, us, "C3 mixed, 512-node tiles")
    dus = np.linalg.norm(us - um) / np.linalg.norm(um)
    print("C3 512-node tiles: %d iterations, GPU true residual %.3e, u vs 1024-node tiles %.2e"
          % (its, trs, dus), flush=True)
    assert res < 1e-8 and ident < 1e-7 and dus < 1e-6 and its <= 12


# ---------------------------------------------------------------------------
# C5a: 16 load cases on a 2M-hex adaptive forest (seeded traction patches)
def test_c5a_sixteen_load_cases():
    t = time.time()
    forests, rho, L = workloads.c5a_adaptive(seed=1, nlevels=5)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    F, params = workloads.patch_loads(forests[0], 16, seed=5)
    print("C5a built: %d hex, %d dofs, 16 patches (%.1f s)" % (prob.nelem, prob.ndof,
                                                               time.time() - t), flush=True)
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=400)
    import torch
    Fd = torch.from_numpy(F).cuda()
    t = time.time()
    U = prob.solve_multi(Fd, opts=opts)
    torch.cuda.synchronize()
    t_b = time.time() - t
    its = [i.iters for i in prob.infos]
    assert all(i.converged for i in prob.infos)
    U = U.cpu().numpy()
    print("C5a batched: %.1f ms, iterations %s" % (1e3 * t_b, its), flush=True)
    orc = po.from_forest(forests[0], BCS)
    s = rho ** Q + K0
    worst = 0.0
    for k in range(16):
        res = orc.true_residual(s, F[k], U[k])
        worst = max(worst, res)
    print("C5a oracle true residual, worst of 16 cases: %.3e" % worst, flush=True)
    assert worst < 1e-8
    # a case solved alone is the same case of the batch, bit for bit
    for k in (0, 7):
        u = prob.solve(F[k], opts=opts)
        assert prob.info.iters == its[k]
        assert np.array_equal(u, U[k])


# ---------------------------------------------------------------------------
# C5b: 96 independent cubes of seeded edges (every integer in 8..48), each
# with coarse levels of its own (not nested where the edge has no coarsen()
# level)
def test_c5b_uneven_batch_per_problem_error():
    t = time.time()
    forests, rho, edges = workloads.cube_batch(96, nlevels=4, seed=0, device=0)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    nc = prob.num_components()
    print("C5b built: %d problems, edges %s, %d hex (%.1f s)"
          % (nc, sorted(edges), prob.nelem, time.time() - t), flush=True)
    assert len(set(edges)) > 20 and any(e % 2 for e in edges)
    assert nc == 96
    opts = solver.default_opts(rtol=1e-8, precision=1, max_iters=300)
    u = prob.solve(f, opts=opts)
    info = [prob.component_info(k) for k in range(nc)]
    assert all(c for _, c, _ in info)
    # per-problem oracle residual: the oracle's operator over the whole forest,
    # norms per component (component node ranges are contiguous)
    orc = po.from_forest(forests[0], BCS)
    s = rho ** Q + K0
    fr = np.where(orc.free_mask(), f, 0.0)
    r = fr - orc.apply_free(s, u)
    conn = np.asarray(forests[0].getMeshConn()).reshape(-1, 8)
    blocks = forests[0].getOctants()["block"]
    nb = [workloads.odd_split(e)[0] ** 3 for e in edges]
    bstart = np.concatenate([[0], np.cumsum(nb)])
    comp_of_elem = np.searchsorted(bstart, blocks, side="right") - 1
    node_comp = np.full(orc.nnodes, -1)
    for a in range(8):
        ind = conn[:, a] >= 0
        node_comp[conn[ind, a]] = comp_of_elem[ind]
    dof_comp = np.repeat(node_comp, 3)
    res = np.array([np.linalg.norm(r[dof_comp == k]) / np.linalg.norm(fr[dof_comp == k])
                    for k in range(nc)])
    print("C5b per-problem oracle true residual: max %.3e median %.3e; GPU's own max %.3e; "
          "iterations %d..%d" % (res.max(), np.median(res), max(x for _, _, x in info),
                                 min(i for i, _, _ in info), max(i for i, _, _ in info)),
          flush=True)
    assert res.max() < 1e-8
    # compliance per problem against the oracle's own solve of that problem
    # (the 12 smallest problems, built alone with their seeds)
    order = np.argsort(edges, kind="stable")[:12]
    worst = 0.0
    for k in order:
        fk, rk = workloads.uneven_cube_batch([int(edges[k])], nlevels=1, seed=int(k))
        ok = po.from_forest(fk[0], BCS)
        ok.assemble(rk ** Q + K0)
        fkv = workloads.face_load(fk[0], [(int(o["tag"]), int(o["info"]))
                                          for o in fk[0].getOctsWithName("load")], [0, 0, -1.0])
        uk = ok.solve(fkv, rtol=1e-12)
        ck_ref = float(fkv @ uk)
        sel = dof_comp == k
        ck = float(f[sel] @ u[sel])
        worst = max(worst, abs(ck - ck_ref) / abs(ck_ref))
    print("C5b compliance vs the oracle's solves (12 smallest problems): worst rel err %.2e"
          % worst, flush=True)
    assert worst < 1e-6
