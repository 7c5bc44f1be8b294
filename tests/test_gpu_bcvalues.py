"""Prescribed non-zero values on clamped dofs (TMRBoundaryConditions bc_vals
handed to TACSAssembler::addBCs): the GPU solve against the FP64 oracle's
lifted solve K_ff u_f = f_f - K_fc u_c."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

Q, K0 = 3.0, 1e-6
BCS = [("fixed", 7), ("load", 2)]   # x- clamped, x+ face: v prescribed


def _case(adaptive):
    if adaptive:
        forests, rho, _ = workloads.adaptive(seed=5, base_level=2, max_level=4, nlevels=3)
    else:
        forests, rho, _ = workloads.cube_l4()
    rho = np.random.default_rng(2).uniform(0.2, 1.0, len(rho))
    f0 = forests[0]
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    n = f0.getNodeRange()[1]
    ubc = np.zeros(3 * n)
    nodes = np.asarray(f0.getNodesWithName("load"))
    nodes = nodes[nodes >= 0]
    ubc[3 * nodes + 1] = 0.01
    return forests, rho, prob, ubc


def _oracle_solution(forest, rho, f, ubc):
    """The oracle library holds one assembled matrix at a time: K without
    conditions for K u_bc, then the clamped operator for the solve, then K
    again for the residual check."""
    s = np.array([po.penalty(r, Q, K0) for r in rho])
    orc = po.from_forest(forest, BCS)
    ptr, dc, dw = forest.getDepNodeConn()
    full = po.Oracle(forest.getMeshConn(), ptr, dc, dw, forest.getElementExtents(),
                     np.zeros(orc.nnodes, np.uint8))
    full.assemble(s)
    Kub = full.matvec(ubc)
    mask = ((orc.bcmask[:, None] >> np.arange(3)) & 1).astype(bool).ravel()
    rhs = f - Kub
    rhs[mask] = 0.0
    orc.assemble(s)
    u = orc.solve(rhs, rtol=1e-12) + ubc
    # the lifted solution satisfies K u = f on the free dofs and takes the
    # values on the clamped ones
    full.assemble(s)
    res = full.matvec(u) - f
    assert np.linalg.norm(res[~mask]) < 1e-9 * (np.linalg.norm(Kub) + np.linalg.norm(f))
    assert np.array_equal(u[mask], ubc[mask])
    return u, full, mask


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_prescribed_values_solve_and_gradient(adaptive, precision):
    forests, rho, prob, ubc = _case(adaptive)
    print("bc values adaptive", adaptive, "precision", precision, "ndof", prob.ndof, flush=True)
    f = prob.traction_load("load", [0.0, 0.0, -0.5])
    u_ref, full, mask = _oracle_solution(forests[0], rho, f, ubc)
    prob.set_bc_values(ubc)
    u = prob.solve(f, opts=solver.default_opts(rtol=1e-10, precision=precision))
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    print("u rel err %.3e" % err, flush=True)
    assert err < 1e-7
    assert np.array_equal(u[mask], ubc[mask])
    fm = f.copy()
    fm[mask] = 0.0
    c_ref = float(fm @ u_ref)
    assert abs(prob.compliance() - c_ref) < 1e-7 * abs(c_ref) + 1e-12
    # TMR's compliance branch: -p'(rho) u_e^T K_e u_e with the full u
    g = prob.compliance_grad()
    g_ref = -np.array([po.dpenalty(r, Q) for r in rho]) * full.energy(u_ref)
    gerr = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print("grad rel err %.3e" % gerr, flush=True)
    assert gerr < 1e-6


def test_zero_values_leave_the_solve_unchanged():
    forests, rho, prob, ubc = _case(False)
    f = prob.traction_load("load", [0.0, 0.0, -0.5])
    opts = solver.default_opts(rtol=1e-10, precision=0)
    u0 = prob.solve(f, opts=opts).copy()
    prob.set_bc_values(np.zeros_like(ubc))
    u1 = prob.solve(f, opts=opts)
    assert np.array_equal(u0, u1)
