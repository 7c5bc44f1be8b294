"""The oracle's multigrid (oracle/mg_oracle.c), pinned on the CPU: its
Galerkin operators equal the triple product P~^T A P~ formed with scipy from
the oracle's assembled operator and interpolation; MG-PCG (Chebyshev) and
GMRES(50) with symmetric-SOR MG (TMRTopoProblem's solver) reach the
Jacobi-PCG solution; one V-cycle on a uniform two-level problem matches a
dense restatement of the cycle."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as po
from tmr_amd import workloads

BCS = [("fixed", 7)]


def traction_z(forest, orc):
    """Consistent load of a unit traction -z on the faces named 'load'
    (exact face integral t A / 4 per face corner, hanging corners to their
    parents)."""
    conn = forest.getMeshConn()
    ext = forest.getElementExtents()
    dp, dc, dw = forest.getDepNodeConn()
    f = np.zeros(3 * orc.nnodes)
    for o in forest.getOctsWithName("load"):
        e, face = o["tag"], o["info"]
        ax = face // 2
        area = ext[e][(ax + 1) % 3] * ext[e][(ax + 2) % 3]
        for k in range(8):
            if [(k & 1), (k >> 1) & 1, (k >> 2)][ax] == face % 2:
                g = conn[e][k]
                if g >= 0:
                    f[3 * g + 2] += -0.25 * area
                else:
                    d = -g - 1
                    for j in range(dp[d], dp[d + 1]):
                        f[3 * dc[j] + 2] += -0.25 * area * dw[j]
    return f


@pytest.fixture(scope="module")
def adapt():
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=4)
    orc = po.from_forest(forests[0], BCS)
    orc.assemble(rho ** 3 + 1e-6)
    return forests, rho, orc


def test_galerkin_matches_scipy_triple_product(adapt):
    forests, rho, orc = adapt
    mg = po.Multigrid(orc, forests, BCS)
    rp, cl, v = orc.csr()
    A = sp.csr_matrix((v, cl, rp), shape=(3 * orc.nnodes, 3 * orc.nnodes))
    for l in range(1, len(forests)):
        _, _, (prp, pcl, pw) = po.interpolation((2, 2, 2), forests[l - 1].getOctants(),
                                                forests[l].getOctants())
        nf, nc = len(prp) - 1, int(forests[l].getNodeRange()[1])
        P = sp.kron(sp.csr_matrix((pw, pcl, prp), shape=(nf, nc)), sp.identity(3), format="csr")
        ff = orc.free_mask() if l == 1 else free
        free = ~(((mg.masks[l][:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))
        Pm = sp.diags(ff.astype(float)) @ P @ sp.diags(free.astype(float))
        A = (Pm.T @ A @ Pm + sp.diags((~free).astype(float))).tocsr()
        G = mg.matrix(l)
        err = abs(G - A).max() / abs(A).max()
        assert err < 1e-13, (l, err)


@pytest.mark.parametrize("mode", ["pcg-chebyshev", "gmres-sor"])
def test_multigrid_solves_match_jacobi_pcg(adapt, mode):
    forests, rho, orc = adapt
    f = traction_z(forests[0], orc)
    u_ref = orc.solve(f, rtol=1e-13)
    mg = po.Multigrid(orc, forests, BCS)
    if mode == "pcg-chebyshev":
        mg.setup("chebyshev", degree=2)
        u = mg.pcg(f, rtol=1e-11, maxit=500)
    else:
        mg.setup("sor", omega=1.0)
        u = mg.gmres(f, m=50, rtol=1e-11, maxit=1000)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-9
    s = rho ** 3 + 1e-6
    assert orc.true_residual(s, f, u) < 1e-8


def test_uniform_density_needs_few_iterations():
    """A uniform cube with uniform density: the Galerkin V-cycle is a strong
    preconditioner (iteration count pinned as a regression bound)."""
    forests, rho, L = workloads.cube_l4(nlevels=3)
    orc = po.from_forest(forests[0], BCS)
    orc.assemble(np.full(len(rho), 1.0))
    f = traction_z(forests[0], orc)
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("chebyshev", degree=2)
    mg.pcg(f, rtol=1e-10)
    assert mg.iters <= 14, mg.iters


def test_single_components(adapt):
    """orc_mg_op (the per-component checker of the GPU's multigrid): the
    level operators and transfers equal the matrices they are built from,
    the V-cycle is a symmetric operator (pre- and post-smoothing by the
    same Chebyshev polynomial), and the coarse solve inverts the coarsest
    Galerkin matrix."""
    forests, rho, orc = adapt
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("chebyshev", degree=2)
    rng = np.random.default_rng(4)
    nl = len(forests)
    free = [~(((m[:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1)) for m in mg.masks]
    for l in range(nl - 1):
        x = rng.standard_normal(len(free[l]))
        assert np.allclose(mg.op("apply", l, x), mg.matrix(l) @ x, rtol=1e-13, atol=1e-13)
        _, _, (prp, pcl, pw) = po.interpolation((2, 2, 2), forests[l].getOctants(),
                                                forests[l + 1].getOctants())
        nf, nc = len(prp) - 1, int(forests[l + 1].getNodeRange()[1])
        P = sp.kron(sp.csr_matrix((pw, pcl, prp), shape=(nf, nc)), sp.identity(3), format="csr")
        Pm = sp.diags(free[l].astype(float)) @ P @ sp.diags(free[l + 1].astype(float))
        assert np.allclose(mg.op("restrict", l, x), Pm.T @ x, rtol=1e-14, atol=1e-14)
        xc = rng.standard_normal(len(free[l + 1]))
        assert np.allclose(mg.op("prolong", l, xc), Pm @ xc, rtol=1e-14, atol=1e-14)
    a = np.where(free[0], rng.standard_normal(len(free[0])), 0.0)
    b = np.where(free[0], rng.standard_normal(len(free[0])), 0.0)
    va, vb = mg.op("vcycle", 0, a), mg.op("vcycle", 0, b)
    assert abs(b @ va - a @ vb) < 1e-10 * abs(b @ va)
    lc = nl - 1
    bc = np.where(free[lc], rng.standard_normal(len(free[lc])), 0.0)
    yc = mg.op("coarse", lc, bc)
    assert np.linalg.norm(mg.matrix(lc) @ yc - bc) < 1e-10 * np.linalg.norm(bc)
    with pytest.raises(ValueError):
        mg.op("coarse", 0, bc)


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_coarse_smoothing_is_the_chebyshev_polynomial(adapt, degree):
    """use_coarse_direct_solve = 0 (TMR_CreateTACSMg hands the coarsest level
    its smoother, src/TMR_RefinementTools.cpp:87-102): the coarse op is one
    Chebyshev sweep from zero, whose residual is fixed by the textbook
    identity D^-1 (b - A x_k) = T_k((theta - M) / delta) / T_k(theta / delta)
    D^-1 b, M = D^-1 A, theta and delta the centre and half-width of
    [lower, upper] lambda_max.  The identity is evaluated densely with the
    three-term recurrence of T_k, not the smoother's update formulas."""
    forests, rho, orc = adapt
    lower, upper = 1.0 / 6.0, 1.1
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("chebyshev", degree=degree, lower=lower, upper=upper, coarse_direct=False)
    lc = len(forests) - 1
    A = mg.matrix(lc).toarray()
    free = ~(((mg.masks[lc][:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))
    lam = mg.get_lambda(lc)
    assert lam > 1.0
    rng = np.random.default_rng(11 + degree)
    b = np.where(free, rng.standard_normal(len(free)), 0.0)
    x = mg.op("coarse", lc, b)
    dinv = 1.0 / np.diag(A)
    M = dinv[:, None] * A

    def residual(theta, delta):
        S = (theta * np.eye(len(b)) - M) / delta
        s0 = theta / delta
        Tp, T = np.eye(len(b)), S.copy()          # T_0, T_1 of S
        tp, t = 1.0, s0                           # T_0, T_1 of s0
        for _ in range(degree - 1):
            Tp, T = T, 2.0 * S @ T - Tp
            tp, t = t, 2.0 * s0 * t - tp
        return (T @ (dinv * b)) / t

    theta, delta = 0.5 * (upper + lower) * lam, 0.5 * (upper - lower) * lam
    got = dinv * (b - A @ x)
    scale = np.linalg.norm(dinv * b)
    err = np.linalg.norm(got - residual(theta, delta)) / scale
    assert err < 1e-12, err
    # a 1% wrong interval is caught (the check has teeth)
    assert np.linalg.norm(got - residual(1.01 * theta, delta)) / scale > 1e-5
    # the hierarchy with a smoothed coarsest level still converges to the
    # Jacobi-PCG solution
    f = traction_z(forests[0], orc)
    u = mg.pcg(f, rtol=1e-11, maxit=800)
    u_ref = orc.solve(f, rtol=1e-13)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-8
    mg.setup("chebyshev", degree=2, coarse_direct=True)


@pytest.mark.parametrize("omega", [1.0, 0.7])
def test_ordered_sor_is_the_ssor_formula(adapt, omega):
    """The oracle's symmetric SOR in a given node order (the check of the
    GPU's colour-ordered SOR, TACSGaussSeidel's role in TMR_CreateTACSMg,
    src/TMR_RefinementTools.cpp:93-94) equals the textbook sweeps on the
    permuted system A' = Q A Q^T, split D' + L' + U': forward
    (D' + w L') x1 = w b' - (w U' + (w - 1) D') x0 from x0 = 0, backward
    (D' + w U') x2 = w b' - (w L' + (w - 1) D') x1, by dense triangular
    solves."""
    import scipy.linalg as sl
    forests, rho, orc = adapt
    mg = po.Multigrid(orc, forests, BCS)
    mg.setup("sor", omega=omega, coarse_direct=False)
    rng = np.random.default_rng(int(omega * 10))
    nl = len(forests)
    for l, op in [(nl - 2, "smooth"), (nl - 1, "coarse")]:
        A = mg.matrix(l).toarray()
        nn = A.shape[0] // 3
        order = rng.permutation(nn).astype(np.intc)
        mg.set_order(l, order)
        free = ~(((mg.masks[l][:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))
        b = np.where(free, rng.standard_normal(len(free)), 0.0)
        y = mg.op(op, l, b)
        perm = (3 * order[:, None] + np.arange(3)).reshape(-1)
        Ap, bp = A[np.ix_(perm, perm)], b[perm]
        D = np.diag(np.diag(Ap))
        Lo, Up = np.tril(Ap, -1), np.triu(Ap, 1)
        x1 = sl.solve_triangular(D + omega * Lo, omega * bp, lower=True)
        x2 = sl.solve_triangular(D + omega * Up, omega * bp - (omega * Lo + (omega - 1) * D) @ x1,
                                 lower=False)
        want = np.empty_like(x2)
        want[perm] = x2
        err = np.linalg.norm(y - want) / np.linalg.norm(want)
        assert err < 1e-12, (l, err)
        # the natural order gives another result (the order is honoured)
        mg.set_order(l, None)
        assert np.linalg.norm(mg.op(op, l, b) - want) / np.linalg.norm(want) > 1e-6
    # SOR-MG with a smoothed coarsest level still solves the system
    f = traction_z(forests[0], orc)
    u = mg.pcg(f, rtol=1e-11, maxit=800)
    u_ref = orc.solve(f, rtol=1e-13)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-8
