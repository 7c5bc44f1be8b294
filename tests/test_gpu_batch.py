"""Batched solves on the GPU: several load cases in one PCG (each case's
arithmetic is that of its own solve, so results are bit-identical), and
independent problems of uneven size side by side in one forest (each
component keeps its own PCG scalars and stopping test)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6


@pytest.fixture(scope="module")
def adapt():
    forests, rho, L = workloads.adaptive(seed=4, base_level=2, max_level=4, nlevels=3)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    return forests, rho, prob


def _loads(prob, n):
    rng = np.random.default_rng(2)
    F = np.zeros((n, prob.ndof))
    F[0] = prob.traction_load("load", [0.0, 0.0, -1.0])
    for k in range(1, n):
        F[k] = prob.traction_load("load", rng.standard_normal(3))
    return F


@pytest.mark.parametrize("precision,no_graph", [(1, 0), (0, 0), (1, 1)])
def test_multi_rhs_bit_identical_to_single(adapt, precision, no_graph):
    forests, rho, prob = adapt
    F = _loads(prob, 5)
    opts = solver.default_opts(rtol=1e-10, precision=precision, no_graph=no_graph)
    print("multi-rhs", F.shape, "precision", precision, "no_graph", no_graph, flush=True)
    U = prob.solve_multi(F, opts=opts)
    its = [i.iters for i in prob.infos]
    assert all(i.converged for i in prob.infos)
    for k in range(F.shape[0]):
        u = prob.solve(F[k], opts=opts)
        print("case", k, "iters", its[k], prob.info.iters, flush=True)
        assert prob.info.iters == its[k]
        assert np.array_equal(u, U[k])
        assert prob.true_residual() < 1e-9


def test_multi_rhs_state_is_case_zero(adapt):
    forests, rho, prob = adapt
    F = _loads(prob, 3)
    opts = solver.default_opts(rtol=1e-10)
    U = prob.solve_multi(F, opts=opts)
    assert np.isclose(prob.compliance(), F[0] @ U[0], rtol=1e-12)


def test_components_match_separate_problems():
    dims = [(1, 1, 1), (2, 1, 1), (1, 1, 2)]
    forests, rho = workloads.box_batch(dims, 3, 3)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    assert prob.num_components() == len(dims)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    # each component stops on its own tolerance; the Chebyshev interval comes
    # from the batch's largest eigenvalue, so iterates differ slightly from
    # the separate solves: compare converged solutions
    opts = solver.default_opts(rtol=1e-11)
    print("components", len(dims), "ndof", prob.ndof, flush=True)
    u = prob.solve(f, opts=opts)
    off = 0
    for k, d in enumerate(dims):
        fk, rk = workloads.box_batch([d], 3, 3, seed=k)
        pk = solver.Problem(fk, q=Q, k0=K0, bcs=BCS)
        pk.set_design(rk)
        gk = pk.traction_load("load", [0.0, 0.0, -1.0])
        assert np.array_equal(gk, f[off:off + pk.ndof])
        uk = pk.solve(gk, opts=opts)
        it, conv, rr = prob.component_info(k)
        err = np.linalg.norm(u[off:off + pk.ndof] - uk) / np.linalg.norm(uk)
        print("component", k, "iters", it, "alone", pk.info.iters, "rel res %.2e" % rr,
              "u diff %.2e" % err, flush=True)
        assert conv and rr < 1e-9
        assert err < 1e-6
        off += pk.ndof
    assert off == prob.ndof


def test_uneven_batch_smoke():
    forests, rho = workloads.box_batch([(1, 1, 1), (3, 3, 3), (2, 2, 2), (1, 1, 1)], 3, 4)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    prob.solve(f, opts=solver.default_opts(rtol=1e-8))
    for k in range(prob.num_components()):
        it, conv, rr = prob.component_info(k)
        assert conv and rr < 1e-7


@pytest.mark.parametrize("precision", [1, 0])
def test_uneven_edges_not_nested(precision):
    """Cubes whose edges have no coarsen() level (9, 13, 11 elements: level-0
    blocks only) beside ones that have (8, 12), each with coarse levels of
    its own (13 -> 7 -> 4 -> 2 elements per edge), so that the transfers
    come from physical positions.  Every problem's solution is checked with
    the oracle's FP64 operator on that problem alone (true residual) and
    against the oracle's own solve of it (u, compliance)."""
    edges = [9, 13, 8, 12, 11]
    forests, rho = workloads.uneven_cube_batch(edges, nlevels=4, seed=2)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    assert prob.num_components() == len(edges)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    opts = solver.default_opts(rtol=1e-11, precision=precision, max_iters=200)
    print("uneven cubes", edges, "precision", precision, "ndof", prob.ndof, flush=True)
    u = prob.solve(f, opts=opts)
    off = 0
    for k, n in enumerate(edges):
        fk, rk = workloads.uneven_cube_batch([n], nlevels=1, seed=2 + k)
        ok = po.from_forest(fk[0], BCS)
        sk = rk ** Q + K0
        ok.assemble(sk)
        nd = 3 * ok.nnodes
        fkv = f[off:off + nd]
        uk = ok.solve(fkv, rtol=1e-13)
        ub = u[off:off + nd]
        it, conv, rr = prob.component_info(k)
        res = ok.true_residual(sk, fkv, ub)
        eu = np.linalg.norm(ub - uk) / np.linalg.norm(uk)
        ec = abs(fkv @ ub - fkv @ uk) / abs(fkv @ uk)
        print("edge %d: %d iterations, oracle true residual %.2e, u vs oracle solve %.2e, "
              "compliance %.2e" % (n, it, res, eu, ec), flush=True)
        assert conv and res < 1e-10
        assert eu < 1e-7 and ec < 1e-9
        # the coarse levels that are not nested work as well as nested ones:
        # 13-14 iterations measured for every edge, the nested 8 and 12 too
        assert it <= 20
        off += nd
    assert off == prob.ndof
    # Galerkin coarse operators need nested forests: a clear error
    with pytest.raises(Exception, match="nested"):
        prob.set_multigrid(use_galerkin=True)


def test_autotune_keeps_results(adapt):
    forests, rho, prob = adapt
    F = _loads(prob, 1)
    opts = solver.default_opts(rtol=1e-10)
    u0 = prob.solve(F[0], opts=opts)
    chosen, ms = prob.autotune((256, 512, 1024), reps=2)
    print("autotune", chosen, ms, flush=True)
    assert chosen in (256, 512, 1024)
    u1 = prob.solve(F[0], opts=opts)
    assert np.linalg.norm(u1 - u0) / np.linalg.norm(u0) < 1e-8


@pytest.mark.parametrize("precision,no_graph", [(1, 0), (0, 1)])
def test_iteration_cap_keeps_last_update(adapt, precision, no_graph):
    """Stopped by the iteration cap, x holds every iteration's alpha p (the
    last one is added by that iteration's p update): the true residual
    |f - K x| / |f| matches the PCG recurrence residual."""
    forests, rho, prob = adapt
    F = _loads(prob, 2)
    opts = solver.default_opts(rtol=1e-14, precision=precision, no_graph=no_graph, max_iters=3)
    print("iteration cap, precision", precision, "no_graph", no_graph, flush=True)
    prob.solve_multi(F, opts=opts, allow_unconverged=True)
    for info in prob.infos:
        assert info.iters == 3 and not info.converged
    rec = prob.infos[0].rnorm / prob.infos[0].rnorm0
    tru = prob.true_residual()
    print("recurrence %.6e true %.6e" % (rec, tru), flush=True)
    assert abs(tru - rec) < 1e-6 * rec
