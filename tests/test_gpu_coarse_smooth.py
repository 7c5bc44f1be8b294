"""The coarsest level smoothed instead of solved (tmrx_mg_create with
use_coarse_direct_solve = 0): TMR_CreateTACSMg then hands TACSMg the
level's smoother there, applied once (src/TMR_RefinementTools.cpp:87-102).

Galerkin hierarchy: the coarse op (one Chebyshev sweep from zero on the
coarsest P^T A P) and the whole V-cycle against the oracle's multigrid set
up the same way, with the GPU's eigenvalue estimates installed.
Rediscretised hierarchy: the coarse op against the oracle's Chebyshev sweep
on the coarsest operator read back column by column from the GPU (the
operator itself is pinned elsewhere), and full solves that reach the
direct-coarse solution."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6
LOWER, UPPER = 1.0 / 6.0, 1.1


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


@pytest.mark.parametrize("precision", [0, 1])
def test_galerkin_coarse_smoothed(hier, precision):
    forests, rho, orc = hier
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("chebyshev", degree=2, lower=LOWER, upper=UPPER, coarse_direct=False)
    free = [_free(m) for m in mg.masks]
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_galerkin=True, use_coarse_direct_solve=False)
    prob.set_design(rho)
    opts = solver.default_opts(precision=precision, fp32_valu=precision)
    nl = len(forests)
    lc = nl - 1
    for l in range(nl):
        _, lam = prob.mg_op("apply", l, np.zeros(len(free[l])), opts)
        print("level %d: lambda_max(D^-1 A) GPU %.6f oracle %.6f" % (l, lam, mg.get_lambda(l)),
              flush=True)
        assert lam > 0.0 and 0.5 < lam / mg.get_lambda(l) < 2.0
        mg.set_lambda(l, lam)
    rng = np.random.default_rng(23 + precision)
    tol = 1e-12 if precision == 0 else 1e-5
    # the coarsest Galerkin operator is applied on the GPU now
    x = np.where(free[lc], rng.standard_normal(len(free[lc])), 0.0)
    ea = _rel(prob.mg_op("apply", lc, x, opts)[0], mg.op("apply", lc, x))
    b = np.where(free[lc], rng.standard_normal(len(free[lc])), 0.0)
    y, _ = prob.mg_op("coarse", lc, b, opts)
    ec = _rel(y, mg.op("coarse", lc, b))
    b0 = np.where(free[0], rng.standard_normal(len(free[0])), 0.0)
    z, _ = prob.mg_op("vcycle", 0, b0, opts)
    ev = _rel(z, mg.op("vcycle", 0, b0))
    print("precision %d: coarsest apply %.2e, coarse sweep %.2e, V-cycle %.2e"
          % (precision, ea, ec, ev), flush=True)
    assert ea < tol and ec < 10 * tol
    assert ev < (1e-8 if precision == 0 else 1e-4)
    # a direct coarse solve would give another output (the check has teeth)
    assert _rel(y, np.linalg.solve(mg.matrix(lc).toarray(), b)) > 1e-3


def test_rediscretised_coarse_sweep(hier):
    forests, rho, orc = hier
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_multigrid(use_coarse_direct_solve=False)
    prob.set_design(rho)
    opts = solver.default_opts(precision=0)
    lc = len(forests) - 1
    mask = po.bcmask_from_forest(forests[lc], BCS)
    n = 3 * len(mask)
    print("coarsest level: %d dofs" % n, flush=True)
    assert n <= 3000
    # the coarsest rediscretised operator, column by column
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j], lam = prob.mg_op("apply", lc, e, opts)
        e[j] = 0.0
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()

    class Coarsest:
        """The oracle's multigrid input for a one-level hierarchy: the
        coarsest operator read back from the GPU."""
        nnodes = len(mask)

        def csr(self):
            import scipy.sparse as sp
            S = sp.csr_matrix(A)
            return (S.indptr.astype(np.intc), S.indices.astype(np.intc), S.data.copy())

    mg = po.Multigrid(Coarsest(), [forests[lc]], BCS)
    mg.setup("chebyshev", degree=2, lower=LOWER, upper=UPPER, coarse_direct=False)
    mg.set_lambda(0, lam)
    rng = np.random.default_rng(29)
    b = np.where(_free(mask), rng.standard_normal(n), 0.0)
    y, _ = prob.mg_op("coarse", lc, b, opts)
    ec = _rel(y, mg.op("coarse", 0, b))
    print("coarse sweep vs oracle %.2e (lambda %.6f)" % (ec, lam), flush=True)
    assert ec < 1e-12


@pytest.mark.parametrize("galerkin", [False, True])
def test_solve_with_smoothed_coarsest(hier, galerkin):
    forests, rho, orc = hier
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    f = prob.traction_load("load", (0.0, 0.0, -1.0))
    prob.set_multigrid(use_galerkin=galerkin)
    prob.set_design(rho)
    opts = solver.default_opts(precision=1, rtol=1e-10, max_iters=400)
    u_d = prob.solve(f, opts=opts)
    it_d = prob.info.iters
    prob.set_multigrid(use_galerkin=galerkin, use_coarse_direct_solve=False)
    u_s = prob.solve(f, opts=opts)
    it_s = prob.info.iters
    res = prob.true_residual()
    d = _rel(u_s, u_d)
    print("galerkin %d: iterations direct coarse %d, smoothed coarse %d; true residual %.2e, "
          "solution difference %.2e" % (galerkin, it_d, it_s, res, d), flush=True)
    assert prob.info.converged and res < 1e-8 and d < 1e-7
    # the oracle's check of the solution, element by element
    assert orc.true_residual(rho ** Q + K0, f, u_s) < 1e-8
    # back to the direct solve: the same iterations as before
    prob.set_multigrid(use_galerkin=galerkin)
    prob.solve(f, opts=opts)
    assert prob.info.iters == it_d
