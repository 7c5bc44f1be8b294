"""GPU parity of the hex8 apply, the MG-PCG solve, compliance and its
gradient against the FP64 CPU oracle (oracle/fe_oracle.c)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6


def _oracle(forest, rho):
    orc = po.from_forest(forest, BCS)
    orc.assemble(np.array([po.penalty(r, Q, K0) for r in rho]) if len(rho) < 200000 else
                 rho ** Q + K0)
    return orc


def _load(prob):
    return prob.traction_load("load", [0.0, 0.0, -1.0])


@pytest.fixture(scope="module")
def cube():
    forests, rho, L = workloads.cube_l4()
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    return forests, rho, prob


@pytest.fixture(scope="module")
def adapt():
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=3)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    return forests, rho, prob


@pytest.mark.parametrize("which", ["cube", "adapt"])
def test_apply_parity(which, cube, adapt):
    forests, rho, prob = cube if which == "cube" else adapt
    print("apply parity", which, "nelem", prob.nelem, "ndof", prob.ndof, flush=True)
    orc = _oracle(forests[0], rho)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(prob.ndof)
    # identity rows on clamped dofs in both operators
    y_ref = orc.matvec(x)
    y64 = prob.apply(x, precision=0)
    y32 = prob.apply(x, precision=1)   # matrix cores, bf16 hi/lo split products
    y32v = prob.apply(x, precision=2)  # vector ALUs, FP32 products
    e64 = np.linalg.norm(y64 - y_ref) / np.linalg.norm(y_ref)
    e32 = np.linalg.norm(y32 - y_ref) / np.linalg.norm(y_ref)
    e32v = np.linalg.norm(y32v - y_ref) / np.linalg.norm(y_ref)
    print("rel err fp64 %.3e fp32 mfma %.3e fp32 valu %.3e" % (e64, e32, e32v), flush=True)
    assert e64 < 1e-13
    assert e32v < 1e-5
    # hi*hi + hi*lo + lo*hi keeps ~16 significant bits per product: the
    # preconditioner-only operator is held to 1e-4
    assert e32 < 1e-4


@pytest.mark.parametrize("nlevels", [1, 2])
def test_large_coarse_direct_solve(nlevels):
    """A coarsest level of at least 2048 dofs takes the super-block Cholesky
    and explicit inverse (256-wide blocks, the bulk in FP64 GEMMs): with one
    level it is the whole 16 x 8 x 8 box (4,131 dofs), so the coarse solve
    must match the oracle's operator and one PCG step solves the problem;
    with two levels (Galerkin) the coarse solve of the 32 x 16 x 16 box's
    coarse level must match the oracle multigrid's own Cholesky solve."""
    forests, rho, L = workloads.box(level=3 if nlevels == 1 else 4, nlevels=nlevels, seed=2)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    orc = _oracle(forests[0], rho)
    rng = np.random.default_rng(5)
    opts = solver.default_opts(rtol=1e-10, precision=0)
    lc = nlevels - 1
    if nlevels == 1:
        free = orc.free_mask()
        b = np.where(free, rng.standard_normal(prob.ndof), 0.0)
        y, _ = prob.mg_op("coarse", lc, b, opts)
        res = np.linalg.norm(orc.matvec(y) - b) / np.linalg.norm(b)
        print("one level, %d dofs: coarse solve residual in the oracle's operator %.2e"
              % (prob.ndof, res), flush=True)
        assert res < 1e-10
    else:
        prob.set_multigrid(use_galerkin=True)
        mg = po.Multigrid(orc, forests, BCS)
        mg.setup("chebyshev", degree=2, lower=1.0 / 6.0, upper=1.1)
        fc = mg.masks[lc]
        freec = ~(((fc[:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))
        b = np.where(freec, rng.standard_normal(len(freec)), 0.0)
        y, _ = prob.mg_op("coarse", lc, b, opts)
        ec = np.linalg.norm(y - mg.op("coarse", lc, b)) / np.linalg.norm(y)
        print("two levels, coarse %d dofs: against the oracle's Cholesky solve %.2e"
              % (len(b), ec), flush=True)
        assert ec < 1e-10
    f = _load(prob)
    u_ref = orc.solve(f, rtol=1e-13)
    u = prob.solve(f, opts=opts)
    eu = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    print("solve: %d iterations, u vs oracle %.2e" % (prob.info.iters, eu), flush=True)
    assert eu < 1e-8 and prob.true_residual() < 1e-10
    if nlevels == 1:
        assert prob.info.iters <= 2


@pytest.mark.parametrize("which", ["cube", "adapt"])
def test_fp32_matrix_cores_fp32_operands(which, cube, adapt):
    """The FP32 operator on the matrix cores with FP32 operands
    (v_mfma_f32_16x16x4_f32): the product against the oracle's FP64
    operator at the FP32 vector-ALU bound, and a mixed-precision solve
    through it (every epilogue: Chebyshev step, residual, product)."""
    forests, rho, prob = cube if which == "cube" else adapt
    orc = _oracle(forests[0], rho)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(prob.ndof)
    y_ref = orc.matvec(x)
    keep = prob.tuning()[1]
    prob.set_fp32_form(2)
    try:
        assert prob.tuning()[1] == 2
        y = prob.apply(x, precision=1)
        e = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
        f = _load(prob)
        u_ref = orc.solve(f, rtol=1e-13)
        c_ref = float(f @ u_ref)
        u = prob.solve(f, opts=solver.default_opts(rtol=1e-10, precision=1))
        c, tr, it = prob.compliance(), prob.true_residual(), prob.info.iters
    finally:
        prob.set_fp32_form(keep)
    print("%s: FP32-operand matrix-core product rel err %.3e; solve %d iterations, true res "
          "%.3e, compliance rel err %.2e" % (which, e, it, tr, abs(c - c_ref) / abs(c_ref)),
          flush=True)
    assert e < 1e-5
    assert tr < 1e-8 and abs(c - c_ref) / abs(c_ref) < 1e-6
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-6


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("which", ["cube", "adapt"])
def test_solve_compliance_parity(which, precision, cube, adapt):
    """precision 0: FP64 V-cycle; 1: FP32 V-cycle on the matrix cores;
    2: FP32 V-cycle on the vector ALUs."""
    forests, rho, prob = cube if which == "cube" else adapt
    print("solve parity", which, "precision", precision, flush=True)
    f = _load(prob)
    orc = _oracle(forests[0], rho)
    u_ref = orc.solve(f, rtol=1e-13)
    c_ref = float(f @ u_ref)
    opts = solver.default_opts(rtol=1e-10, precision=min(precision, 1),
                               fp32_valu=int(precision == 2))
    u = prob.solve(f, opts=opts)
    c = prob.compliance()
    tr = prob.true_residual()
    print("iters", prob.info.iters, "true res %.3e" % tr, "c %.12g ref %.12g" % (c, c_ref), flush=True)
    assert prob.info.converged
    assert tr < 1e-8
    assert abs(c - c_ref) / abs(c_ref) < 1e-6
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-6


def test_compliance_gradient_parity(cube):
    forests, rho, prob = cube
    f = _load(prob)
    prob.solve(f, opts=solver.default_opts(rtol=1e-11, precision=1))
    g = prob.compliance_grad()
    orc = _oracle(forests[0], rho)
    u_ref = orc.solve(f, rtol=1e-13)
    e = orc.energy(u_ref)
    g_ref = -np.array([po.dpenalty(r, Q) for r in rho]) * e
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("grad rel err %.3e" % err, flush=True)
    assert err < 1e-6


def test_uniform_density_invariant(cube):
    """c(rho) = c(1) (1 + k0) / (rho^q + k0) for a uniform density."""
    forests, rho, prob = cube
    f = _load(prob)
    opts = solver.default_opts(rtol=1e-11, precision=1)
    prob.set_design(np.full(prob.nelem, 0.5))
    prob.solve(f, opts=opts)
    c_half = prob.compliance()
    prob.set_design(np.full(prob.nelem, 1.0))
    prob.solve(f, opts=opts)
    c_one = prob.compliance()
    prob.set_design(rho)
    ratio = c_half / (c_one * (1 + K0) / (0.5 ** Q + K0))
    assert abs(ratio - 1) < 1e-8


@pytest.mark.parametrize("pre,post", [(1, 2), (1, 3), (3, 2)])
def test_asymmetric_cycles_converge(pre, post, cube):
    """Pre- and post-smoothing degrees differ: the first smoothing step lands
    in the buffer that the cycle's parity requires (an unsymmetric cycle is
    a weaker preconditioner, hence the well-conditioned cube)."""
    forests, rho, prob = cube
    f = _load(prob)
    orc = _oracle(forests[0], rho)
    c_ref = float(f @ orc.solve(f, rtol=1e-13))
    opts = solver.default_opts(rtol=1e-10, precision=1, cheb_degree=post, cheb_pre_degree=pre,
                               max_iters=500)
    prob.solve(f, opts=opts)
    print("pre %d post %d iters %d" % (pre, post, prob.info.iters), flush=True)
    assert prob.true_residual() < 1e-9
    assert abs(prob.compliance() - c_ref) / abs(c_ref) < 1e-6


@pytest.mark.parametrize("level", [2, 3])
def test_single_level_is_a_direct_solve(level):
    """With one level the preconditioner is the coarsest-level solve alone:
    the explicit inverse from the GPU's blocked Cholesky (375 dofs in 6
    blocks of 64 at level 2, 2,187 dofs in 35 blocks at level 3, padding
    included), so PCG converges in one iteration to the oracle's solution."""
    from tmr_amd import TMR
    f0 = TMR.createBoxForest(1, 1, 1, (1.0, 1.0, 1.0), face_names=workloads.FACES)
    f0.createTrees(level)
    forests = workloads.hierarchy(f0, 1)
    rho = np.random.default_rng(level).uniform(0.1, 1.0, len(forests[0].getOctants()))
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS)
    prob.set_design(rho)
    f = _load(prob)
    u = prob.solve(f, opts=solver.default_opts(rtol=1e-12, precision=0))
    orc = _oracle(forests[0], rho)
    u_ref = orc.solve(f, rtol=1e-14)
    print("level %d: %d dofs, %d iterations, true res %.2e" % (level, prob.ndof, prob.info.iters,
          prob.true_residual()), flush=True)
    assert prob.info.iters <= 2
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-10
