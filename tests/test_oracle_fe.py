"""Pins of the FE oracle (oracle/fe_oracle.c) from closed forms, invariants,
a dense direct solve and central differences; and the product's host
element matrix (tmr_amd/csrc/fe.cpp via the GPU-free oracle comparison in
test_gpu_solver) is checked against the same oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po

E, NU = 1.0, 0.3
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))


def test_unit_cube_closed_form_entries():
    K = po.hex8_stiffness(1.0, 1.0, 1.0, E, NU)
    # K[ux0, ux0] = (lam + 2 mu) * 1/9 + mu * 2/9
    assert K[0, 0] == pytest.approx((LAM + 4 * MU) / 9.0, rel=1e-14)
    # K[ux0, uy0] = (lam + mu) * (1/2)(1/2)(1/3)
    assert K[0, 1] == pytest.approx((LAM + MU) / 12.0, rel=1e-14)
    # K[ux0, ux7] (opposite corners): int dN0/dx dN7/dx = -1/36 ... etc
    g = np.array([-1 / 36.0, -1 / 36.0, -1 / 36.0])
    assert K[0, 21] == pytest.approx((LAM + 2 * MU) * g[0] + MU * (g[1] + g[2]), rel=1e-13)


def test_symmetry_rank_and_rigid_modes():
    K = po.hex8_stiffness(1.0, 0.5, 0.25, 2.0, 0.3)
    assert np.allclose(K, K.T, atol=1e-14)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12 * ev.max()) == 6
    X = np.array([[(a & 1) * 1.0, ((a >> 1) & 1) * 0.5, (a >> 2) * 0.25] for a in range(8)])
    modes = []
    for d in range(3):
        t = np.zeros((8, 3)); t[:, d] = 1; modes.append(t.ravel())
    for (i, j) in [(0, 1), (1, 2), (0, 2)]:
        r = np.zeros((8, 3)); r[:, i] = -X[:, j]; r[:, j] = X[:, i]; modes.append(r.ravel())
    for m in modes:
        assert np.linalg.norm(K @ m) < 1e-13


def test_constant_strain_patch_energy():
    hx, hy, hz = 0.7, 1.3, 0.4
    K = po.hex8_stiffness(hx, hy, hz, E, NU)
    G = np.array([[0.1, 0.02, -0.03], [0.05, -0.2, 0.01], [0.0, 0.04, 0.15]])
    X = np.array([[(a & 1) * hx, ((a >> 1) & 1) * hy, (a >> 2) * hz] for a in range(8)])
    u = (X @ G.T).ravel()
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0], G[0, 1] + G[1, 0]])
    C = np.zeros((6, 6)); C[:3, :3] = LAM; C[np.arange(3), np.arange(3)] += 2 * MU
    C[3, 3] = C[4, 4] = C[5, 5] = MU
    assert u @ K @ u == pytest.approx(hx * hy * hz * eps @ C @ eps, rel=1e-13)


def test_penalties():
    assert po.penalty(0.5, 3.0, 1e-6, 0) == pytest.approx(0.125 + 1e-6, rel=1e-15)
    assert po.dpenalty(0.5, 3.0, 0) == pytest.approx(0.75, rel=1e-15)
    q = 5.0
    assert po.penalty(0.5, q, 0.0, 1) == pytest.approx(0.5 / (1 + q * 0.5), rel=1e-15)
    h = 1e-6
    fd = (po.penalty(0.5 + h, q, 0.0, 1) - po.penalty(0.5 - h, q, 0.0, 1)) / (2 * h)
    assert po.dpenalty(0.5, q, 1) == pytest.approx(fd, rel=1e-8)


def _small_problem(level=2, dims=(2, 1, 1), seed=0, adaptive=False):
    from tmr_amd import TMR
    f = TMR.createBoxForest(*dims, L=tuple(float(d) for d in dims),
                            face_names={"x-": "fixed", "x+": "load"})
    f.createTrees(level)
    if adaptive:
        rng = np.random.default_rng(seed)
        f.refine(rng.integers(0, 2, len(f.getOctants())).astype(np.intc), 0, level + 2)
    f.balance(1)
    f.createNodes()
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 1.0, len(f.getOctants()))
    orc = po.from_forest(f, [("fixed", 7)])
    s = np.array([po.penalty(r, 3.0, 1e-6) for r in rho])
    orc.assemble(s)
    b = rng.standard_normal(3 * orc.nnodes)
    mask = ((orc.bcmask[:, None] >> np.arange(3)) & 1).astype(bool).ravel()
    b[mask] = 0.0
    return f, orc, rho, s, b


@pytest.mark.parametrize("adaptive", [False, True])
def test_cg_matches_dense_cholesky(adaptive):
    f, orc, rho, s, b = _small_problem(adaptive=adaptive)
    rowp, cols, vals = orc.csr()
    n = len(rowp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, cols[rowp[i]:rowp[i + 1]]] = vals[rowp[i]:rowp[i + 1]]
    assert np.allclose(A, A.T, atol=1e-14)
    L = np.linalg.cholesky(A)
    x_ref = np.linalg.solve(L.T, np.linalg.solve(L, b))
    x = orc.solve(b, rtol=1e-14)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10
    # energy identity b^T x = x^T A x
    assert b @ x == pytest.approx(x @ (A @ x), rel=1e-10)


def test_compliance_gradient_central_differences():
    f, orc, rho, s, b = _small_problem(level=1, seed=3)
    b[:] = 0
    # unit traction on x+ face nodes (z component)
    for n in f.getNodesWithName("load"):
        if n >= 0:
            b[3 * n + 2] = -1.0
    u = orc.solve(b, rtol=1e-14)
    e = orc.energy(u)
    g = -np.array([po.dpenalty(r, 3.0) for r in rho]) * e
    h = 1e-6
    rng = np.random.default_rng(1)
    for k in rng.choice(len(rho), 6, replace=False):
        cs = []
        for sgn in (+1, -1):
            r2 = rho.copy(); r2[k] += sgn * h
            orc.assemble(np.array([po.penalty(r, 3.0, 1e-6) for r in r2]))
            cs.append(b @ orc.solve(b, rtol=1e-14))
        fd = (cs[0] - cs[1]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(cs[0]))


def test_von_mises_closed_form_uniaxial():
    """Uniaxial strain state: vm of C eps by hand."""
    f, orc, rho, s, b = _small_problem(level=1, seed=2)
    # u = strain field eps_xx = 1e-3 everywhere (linear in x), no clamps used
    orc2 = po.Oracle(orc.conn, orc.dep_ptr, orc.dep_conn, orc.dep_w, orc.hdim,
                     np.zeros_like(orc.bcmask), E, NU)
    X = f.getPoints()
    nd = len(f.getDepNodeConn()[0]) - 1
    u = np.zeros(3 * orc.nnodes)
    for n in range(orc.nnodes):
        u[3 * n] = 1e-3 * X[n + nd, 0]
    rho1 = np.ones(orc.nelem)
    val = po.stress_pnorm(orc2, rho1, u, P=8.0, ys=1.0, eps=0.3)
    s0, s1 = (LAM + 2 * MU) * 1e-3, LAM * 1e-3
    vm = abs(s0 - s1)  # sqrt(0.5*(2 (s0-s1)^2)) = |s0 - s1|
    vol = orc.hdim.reshape(-1, 3).prod(1).sum()
    assert val == pytest.approx(vm * vol ** (1 / 8.0), rel=1e-12)


def test_stress_gradient_central_differences():
    f, orc, rho, s, b = _small_problem(level=1, seed=4)
    b[:] = 0
    for n in f.getNodesWithName("load"):
        if n >= 0:
            b[3 * n + 2] = -1.0
    val, g, u, psi = po.stress_gradient(orc, rho, b)
    h = 1e-6
    rng = np.random.default_rng(2)
    for k in rng.choice(len(rho), 6, replace=False):
        vals = []
        for sgn in (+1, -1):
            r2 = rho.copy(); r2[k] += sgn * h
            orc.assemble(np.array([po.penalty(r, 3.0, 1e-6) for r in r2]))
            u2 = orc.solve(b, rtol=1e-14)
            vals.append(po.stress_pnorm(orc, r2, u2))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-8 * abs(val))


@pytest.mark.parametrize("adaptive", [False, True])
def test_element_by_element_product_equals_assembled(adaptive):
    """orc_apply_free (the full-size residual check) against the assembled
    CSR product, clamped rows and columns included."""
    f, orc, rho, s, b = _small_problem(adaptive=adaptive)
    x = np.random.default_rng(5).standard_normal(3 * orc.nnodes)
    y_ref = orc.matvec(x)
    y = orc.apply_free(s, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 1e-14
    fr = np.where(orc.free_mask(), b, 0.0)
    u = orc.solve(fr, rtol=1e-13)
    assert orc.true_residual(s, fr, u) < 1e-12
