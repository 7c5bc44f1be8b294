"""Pins of the oracle's nodal-design restatement (per-Gauss-point penalty,
weighted design restriction, nodal gradients) by checks that do not retype
its formulas: an independent high-order quadrature, a known answer of the
weighted restriction, and central differences."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _iso_C(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def _B(h, xi):
    """6 x 24 strain-displacement matrix of a box element at xi in [-1,1]^3
    (engineering shear order yz, xz, xy), written from the shape functions
    (1 + s_d xi_d)/2 directly."""
    B = np.zeros((6, 24))
    for a in range(8):
        s = np.array([1.0 if (a >> d) & 1 else -1.0 for d in range(3)])
        f = 0.5 * (1 + s * xi)
        dN = np.array([s[0] / h[0] * f[1] * f[2], f[0] * s[1] / h[1] * f[2],
                       f[0] * f[1] * s[2] / h[2]])
        B[0, 3 * a] = dN[0]
        B[1, 3 * a + 1] = dN[1]
        B[2, 3 * a + 2] = dN[2]
        B[3, 3 * a + 1], B[3, 3 * a + 2] = dN[2], dN[1]
        B[4, 3 * a], B[4, 3 * a + 2] = dN[2], dN[0]
        B[5, 3 * a], B[5, 3 * a + 1] = dN[1], dN[0]
    return B


def test_gauss_point_stiffness_integrates_a_trilinear_density():
    """sum_q rho(xi_q) K_q == int rho B^T C B dV for a trilinear rho: the
    integrand is cubic per coordinate, so 2-point Gauss is exact; here it is
    compared with 4-point Gauss.  A mismatch between the Gauss-point order of
    the shape-function table and of K_q fails this."""
    h = np.array([1.0, 0.5, 2.0])
    C = _iso_C(1.0, 0.3)
    rng = np.random.default_rng(0)
    x = rng.uniform(0.2, 1.0, 8)  # corner values, corner a = i + 2 j + 4 k
    rho_q = po.gauss_shape() @ x
    K = sum(rho_q[q] * po.hex8_stiffness_gp(*h, q) for q in range(8))
    pts, wts = np.polynomial.legendre.leggauss(4)
    Kref = np.zeros((24, 24))
    detJ = np.prod(h) / 8
    for i, wi in zip(pts, wts):
        for j, wj in zip(pts, wts):
            for k, wk in zip(pts, wts):
                xi = np.array([i, j, k])
                N = np.array([np.prod([0.5 * (1 + (1 if (a >> d) & 1 else -1) * xi[d])
                                       for d in range(3)]) for a in range(8)])
                B = _B(h, xi)
                Kref += wi * wj * wk * detJ * (N @ x) * B.T @ C @ B
    assert np.max(np.abs(K - Kref)) < 1e-13 * np.max(np.abs(Kref))
    for q in range(8):
        Kq = po.hex8_stiffness_gp(*h, q)
        assert np.allclose(Kq, Kq.T, atol=1e-15)
        ev = np.linalg.eigvalsh(Kq)
        assert ev.min() > -1e-12 * ev.max()
        assert np.sum(ev > 1e-10 * ev.max()) == 6  # one point: rank of B is 6


def _mesh(level=2, seed=None, dims=(2, 1, 1)):
    from tmr_amd import TMR
    f = TMR.createBoxForest(*dims, L=tuple(float(d) for d in dims),
                            face_names={"x-": "fixed", "x+": "load"})
    f.createTrees(level)
    if seed is not None:
        rng = np.random.default_rng(seed)
        f.refine(rng.integers(0, 2, len(f.getOctants())).astype(np.intc), 0, level + 1)
    f.balance(1)
    f.createNodes()
    return f


def test_uniform_nodal_design_equals_element_design():
    f = _mesh(level=1, seed=4)
    orc = po.from_forest(f, [("fixed", 7)])
    assert len(f.getDepNodeConn()[0]) > 1  # hanging nodes present
    x = np.full(orc.nnodes, 0.7)
    rg = orc.gp_density(x)
    assert np.allclose(rg, 0.7, rtol=0, atol=1e-15)
    orc.assemble_gp(np.full(8 * orc.nelem, 0.343))
    A1 = orc.csr()
    orc.assemble(np.full(orc.nelem, 0.343))
    A2 = orc.csr()
    assert np.array_equal(A1[0], A2[0]) and np.array_equal(A1[1], A2[1])
    assert np.max(np.abs(A1[2] - A2[2])) < 1e-14 * np.max(np.abs(A2[2]))


def test_weight_transpose_known_answer():
    """One cube refined once: a coarse corner collects itself (1), three edge
    midpoints (1/2), three face centres (1/4) and the centre (1/8); the
    normalised transpose of x = first coordinate is 0.5625/3.375 = 1/6 at the
    corners on x = 0 and 5/6 on x = 1; constants are kept."""
    H = 1 << 30
    fine = np.array([(0, i * H // 2, j * H // 2, k * H // 2, 0, 1, 0)
                     for k in range(2) for j in range(2) for i in range(2)], dtype=po.OCT_DT)
    coarse = np.array([(0, 0, 0, 0, 0, 0, 0)], dtype=po.OCT_DT)
    fm, cm, (rowp, cols, w) = po.interpolation((1, 1, 1), fine, coarse)
    assert fm.nindep == 27 and cm.nindep == 8
    xf = fm.gxyz[:fm.nindep, 0] / float(H)
    xc = po.weight_transpose(rowp, cols, w, xf, cm.nindep)
    cx = cm.gxyz[:cm.nindep, 0] / float(H)
    assert np.allclose(xc[cx == 0.0], 1.0 / 6.0, rtol=0, atol=1e-15)
    assert np.allclose(xc[cx == 1.0], 5.0 / 6.0, rtol=0, atol=1e-15)
    ones = po.weight_transpose(rowp, cols, w, np.ones(27), 8)
    assert np.array_equal(ones, np.ones(8))
    r = np.random.default_rng(1).uniform(-1, 2, 27)
    rc = po.weight_transpose(rowp, cols, w, r, 8)
    assert np.all(rc >= r.min()) and np.all(rc <= r.max())


def _loads(f, orc):
    b = np.zeros(3 * orc.nnodes)
    for n in f.getNodesWithName("load"):
        if n >= 0:
            b[3 * n + 2] = -1.0
    return b


@pytest.mark.parametrize("ptype", [0, 1])
def test_nodal_compliance_gradient_central_differences(ptype):
    f = _mesh(level=1, seed=5)
    orc = po.from_forest(f, [("fixed", 7)])
    b = _loads(f, orc)
    rng = np.random.default_rng(2)
    x = rng.uniform(0.3, 1.0, orc.nnodes)
    q = 3.0 if ptype == 0 else 5.0
    c, g, _ = po.nodal_compliance_gradient(orc, x, b, q=q, ptype=ptype)
    # parents of hanging nodes take part through the dependent weights
    ptr, dc, dw = f.getDepNodeConn()
    nodes = list(rng.choice(orc.nnodes, 4, replace=False)) + [int(dc[0]), int(dc[-1])]
    h = 1e-6
    for j in nodes:
        cs = []
        for sgn in (1, -1):
            x2 = x.copy()
            x2[j] += sgn * h
            cs.append(po.nodal_compliance_gradient(orc, x2, b, q=q, ptype=ptype)[0])
        fd = (cs[0] - cs[1]) / (2 * h)
        assert g[j] == pytest.approx(fd, rel=2e-5, abs=1e-8 * abs(c))


def test_nodal_stress_gradient_directional_difference():
    f = _mesh(level=1, seed=6)
    orc = po.from_forest(f, [("fixed", 7)])
    b = _loads(f, orc)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.3, 1.0, orc.nnodes)
    fval, g, _, _ = po.nodal_stress_gradient(orc, x, b, P=6.0)
    d = rng.standard_normal(orc.nnodes)
    h = 1e-6
    fp = po.nodal_stress_gradient(orc, x + h * d, b, P=6.0)[0]
    fm = po.nodal_stress_gradient(orc, x - h * d, b, P=6.0)[0]
    assert g @ d == pytest.approx((fp - fm) / (2 * h), rel=1e-5)
