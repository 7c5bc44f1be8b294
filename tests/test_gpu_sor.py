"""Symmetric SOR smoothing (tmrx_mg_create with use_chebyshev_smoother = 0):
TMR_CreateTACSMg's default smoother, TACSGaussSeidel with one symmetric
sweep and relaxation omega (src/TMR_RefinementTools.cpp:45-52, 93-94).

The GPU sweeps the dofs of every smoothed level in colour-major node order
(tmrx_problem_smoother_order); the oracle's sequential SOR is run in that
order (its ordered sweeps are pinned against the textbook SSOR formula on
the CPU) and must give the same smoothing, level by level and through the
whole V-cycle.  Full solves with SOR-MG reach the oracle's solution."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6


@pytest.fixture(scope="module")
def hier():
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=4)
    orc = po.from_forest(forests[0], BCS)
    orc.assemble(rho ** Q + K0)
    print("hierarchy: %s hex" % [len(f.getOctants()) for f in forests], flush=True)
    return forests, rho, orc


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _free(mask):
    return ~(((mask[:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))


def _check_colouring(A, nodes, cptr):
    """No two nodes of one colour share a block of A (dof-level CSR)."""
    colour = np.empty(len(nodes), np.intc)
    for c in range(len(cptr) - 1):
        colour[nodes[cptr[c]:cptr[c + 1]]] = c
    assert np.array_equal(np.sort(nodes), np.arange(len(nodes)))
    C = A.tocoo()
    i, j = C.row // 3, C.col // 3
    off = (i != j) & (C.data != 0.0)
    assert not np.any(colour[i[off]] == colour[j[off]])


@pytest.mark.parametrize("precision,omega", [(0, 1.0), (0, 0.8), (1, 1.0)])
def test_sor_components_galerkin(hier, precision, omega):
    forests, rho, orc = hier
    nl = len(forests)
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("sor", omega=omega)
    free = [_free(m) for m in mg.masks]
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_galerkin=True, use_chebyshev_smoother=False, omega=omega)
    prob.set_design(rho)
    opts = solver.default_opts(precision=precision)
    prob.mg_op("apply", 0, np.zeros(len(free[0])), opts)  # setup
    rng = np.random.default_rng(31 + precision)
    tol = 1e-12 if precision == 0 else 1e-5
    for l in range(nl - 1):
        nodes, cptr = prob.smoother_order(l)
        _check_colouring(mg.matrix(l), nodes, cptr)
        mg.set_order(l, nodes)
        b = np.where(free[l], rng.standard_normal(len(free[l])), 0.0)
        y, _ = prob.mg_op("smooth", l, b, opts)
        es = _rel(y, mg.op("smooth", l, b))
        print("level %d (precision %d, omega %.1f): %d colours, SOR sweep vs oracle %.2e"
              % (l, precision, omega, len(cptr) - 1, es), flush=True)
        assert es < tol
    # the order matters: the oracle's row order gives another sweep
    mg.set_order(0, None)
    b = np.where(free[0], rng.standard_normal(len(free[0])), 0.0)
    y, _ = prob.mg_op("smooth", 0, b, opts)
    assert _rel(y, mg.op("smooth", 0, b)) > 1e-4
    nodes, _ = prob.smoother_order(0)
    mg.set_order(0, nodes)
    z, _ = prob.mg_op("vcycle", 0, b, opts)
    ev = _rel(z, mg.op("vcycle", 0, b))
    print("V-cycle (precision %d, omega %.1f) vs oracle %.2e" % (precision, omega, ev), flush=True)
    assert ev < (1e-10 if precision == 0 else 1e-4)


def test_sor_rediscretised_finest(hier):
    """Rediscretised coarse levels: the finest level's matrix is the same,
    so is its SOR sweep."""
    forests, rho, orc = hier
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("sor", omega=1.0)
    free = _free(mg.masks[0])
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_chebyshev_smoother=False)
    prob.set_design(rho)
    opts = solver.default_opts(precision=0)
    prob.mg_op("apply", 0, np.zeros(len(free)), opts)
    nodes, cptr = prob.smoother_order(0)
    mg.set_order(0, nodes)
    b = np.where(free, np.random.default_rng(5).standard_normal(len(free)), 0.0)
    es = _rel(prob.mg_op("smooth", 0, b, opts)[0], mg.op("smooth", 0, b))
    print("rediscretised finest SOR sweep vs oracle %.2e (%d colours)" % (es, len(cptr) - 1),
          flush=True)
    assert es < 1e-12
    with pytest.raises(Exception):
        prob.smoother_order(len(forests) - 1)  # the coarsest level is solved directly


@pytest.mark.parametrize("galerkin", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_sor_mg_solves(hier, galerkin, precision):
    forests, rho, orc = hier
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_galerkin=galerkin, use_chebyshev_smoother=False)
    prob.set_design(rho)
    f = prob.traction_load("load", (0.0, 0.0, -1.0))
    opts = solver.default_opts(precision=precision, rtol=1e-10, max_iters=400)
    u = prob.solve(f, opts=opts)
    it = prob.info.iters
    u_ref = orc.solve(f, rtol=1e-13)
    d = _rel(u, u_ref)
    res = orc.true_residual(rho ** Q + K0, f, u)
    # Chebyshev smoothing on the same hierarchy, for the record
    prob.set_multigrid(use_galerkin=galerkin)
    prob.solve(f, opts=opts)
    it_c = prob.info.iters
    print("galerkin %d precision %d: SOR-MG %d iterations, Chebyshev %d; vs oracle %.2e, "
          "true residual %.2e" % (galerkin, precision, it, it_c, d, res), flush=True)
    assert res < 1e-8 and d < 1e-7
    # two load cases in one batched solve: each equals its own solve
    prob.set_multigrid(use_galerkin=galerkin, use_chebyshev_smoother=False)
    F = np.stack([f, np.roll(f, 3)])
    U = prob.solve_multi(F, opts=opts)
    u2 = prob.solve(F[1], opts=opts)
    assert _rel(U[0], u) < 1e-12 and _rel(U[1], u2) < 1e-12


def test_sor_coarse_smoothed_galerkin(hier):
    """SOR on every level, the coarsest included (use_coarse_direct_solve =
    0: TMR_CreateTACSMg gives that level its TACSGaussSeidel,
    src/TMR_RefinementTools.cpp:87-102)."""
    forests, rho, orc = hier
    nl = len(forests)
    lc = nl - 1
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("sor", omega=1.0, coarse_direct=False)
    free = [_free(m) for m in mg.masks]
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_galerkin=True, use_coarse_direct_solve=False,
                       use_chebyshev_smoother=False)
    prob.set_design(rho)
    opts = solver.default_opts(precision=0)
    prob.mg_op("apply", 0, np.zeros(len(free[0])), opts)
    for l in range(nl):
        nodes, cptr = prob.smoother_order(l)
        _check_colouring(mg.matrix(l), nodes, cptr)
        mg.set_order(l, nodes)
    rng = np.random.default_rng(41)
    b = np.where(free[lc], rng.standard_normal(len(free[lc])), 0.0)
    ec = _rel(prob.mg_op("coarse", lc, b, opts)[0], mg.op("coarse", lc, b))
    b0 = np.where(free[0], rng.standard_normal(len(free[0])), 0.0)
    ev = _rel(prob.mg_op("vcycle", 0, b0, opts)[0], mg.op("vcycle", 0, b0))
    print("SOR on the coarsest level: coarse op vs oracle %.2e, V-cycle %.2e" % (ec, ev),
          flush=True)
    assert ec < 1e-12 and ev < 1e-10


def _nodal_x(forest, seed):
    """A smooth seeded nodal field in [0.05, 1] on the independent nodes."""
    nums = np.asarray(forest.getNodeNumbers())
    X = forest.getPoints()[nums >= 0]
    rng = np.random.default_rng(seed)
    k = rng.uniform(1.0, 3.0, 3)
    v = 0.55 + 0.45 * np.sin(k[0] * X[:, 0]) * np.cos(k[1] * X[:, 1] + k[2] * X[:, 2])
    return np.clip(v, 0.05, 1.0)


@pytest.mark.parametrize("galerkin", [False, True])
def test_sor_nodal_design(galerkin):
    """Nodal design variables (per-Gauss-point stiffness): the assembled
    level matrices SOR smooths come from the Gauss-point element matrices;
    the SOR-MG solve reaches the oracle's nodal solve."""
    forests, _, _ = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=3)
    x = _nodal_x(forests[0], 11)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design_nodal(x)
    prob.set_multigrid(use_galerkin=galerkin, use_chebyshev_smoother=False)
    f = prob.traction_load("load", (0.0, 0.0, -1.0))
    orc = po.from_forest(forests[0], BCS)
    po.nodal_assemble(orc, x, Q, K0, 0)
    u_ref = orc.solve(f, rtol=1e-13)
    for precision in (0, 1):
        u = prob.solve(f, opts=solver.default_opts(precision=precision, rtol=1e-10,
                                                   max_iters=400))
        d = _rel(u, u_ref)
        print("nodal SOR-MG galerkin %d precision %d: %d iterations, vs oracle %.2e"
              % (galerkin, precision, prob.info.iters, d), flush=True)
        assert prob.info.converged and d < 1e-7
