"""Helmholtz and matrix design filters on the GPU (tmrx_filter_*) against
the oracle's restatement on assembled operators whose element matrices are
tensor products of 1D linear elements (oracle/pyoracle.py), closed forms
(a constant field passes both filters unchanged: K 1 = 0, and the matrix
filter is normalised by its response to 1), the adjoint identity
<F x, y> = <x, F^T y>, and a TopoProblem with each filter checked by
central differences."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import TMR, TopOptUtils, topo, workloads

pytestmark = pytest.mark.gpu

FACES = {"x-": "fixed", "x+": "load"}


@pytest.fixture(scope="module")
def adapt():
    forests, rho, L = workloads.adaptive(seed=8, base_level=2, max_level=4, nlevels=2)
    M, K = po.scalar_mass_laplacian(forests[0])
    return forests[0], M, K


def _native(forest, kind, r, N=0):
    import ctypes as C
    from tmr_amd._lib import check, lib
    h = C.c_void_p()
    check(lib().tmrx_filter_create(forest.ptr, kind, r, N, 0, C.byref(h)), "filter")

    def run(x, transpose=False, nv=1):
        x = np.ascontiguousarray(x, dtype=np.double)
        out = np.zeros_like(x)
        fn = lib().tmrx_filter_apply_transpose if transpose else lib().tmrx_filter_apply
        check(fn(h, nv, C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data), 0), "apply")
        return out

    def iters():
        it = C.c_int()
        check(lib().tmrx_filter_info(h, None, C.byref(it)))
        return it.value
    return run, iters, lambda: lib().tmrx_filter_free(h)


@pytest.mark.parametrize("kind,r,N", [(0, 0.05, 0), (0, 0.2, 0), (1, 0.05, 10), (1, 0.15, 4)])
def test_filter_matches_oracle(adapt, kind, r, N):
    forest, M, K = adapt
    run, iters, free = _native(forest, kind, r, N)
    n = M.shape[0]
    rng = np.random.default_rng(kind * 10 + N)
    x, y = rng.random(n), rng.standard_normal(n)
    fx, fty = run(x), run(y, transpose=True)
    if kind == 0:
        rx, rty = po.helmholtz_filter(M, K, r, x), po.helmholtz_filter(M, K, r, y, transpose=True)
    else:
        rx, rty = po.matrix_filter(M, r, N, x), po.matrix_filter(M, r, N, y, transpose=True)
    e1 = np.linalg.norm(fx - rx) / np.linalg.norm(rx)
    e2 = np.linalg.norm(fty - rty) / np.linalg.norm(rty)
    print("%s r=%g N=%d: %d nodes, filter rel err %.2e, transpose %.2e, CG its %d"
          % ("helmholtz" if kind == 0 else "matrix", r, N, n, e1, e2, iters()), flush=True)
    assert e1 < 1e-10 and e2 < 1e-10
    # a constant field passes unchanged (up to the Helmholtz filter's CG
    # tolerance, 1e-12 in the residual, times the operator's condition number)
    dev = np.abs(run(np.full(n, 0.7)) - 0.7).max() / 0.7
    print("constant field: max relative deviation %.2e" % dev, flush=True)
    assert dev < 1e-9
    # adjoint identity
    assert abs(fx @ y - x @ fty) <= 1e-11 * abs(fx @ y)
    # interleaved values per node: each component filtered alone
    x2 = np.stack([x, 2.0 * x - 0.3], 1).reshape(-1)
    f2 = run(x2, nv=2).reshape(-1, 2)
    assert np.allclose(f2[:, 0], fx, rtol=1e-12, atol=1e-14)
    assert np.allclose(f2[:, 1], 2.0 * fx - 0.3, rtol=1e-9, atol=1e-12)
    free()


def _mfilter(assemblers, filters):
    mf = TopOptUtils.Mfilter(4, assemblers, filters, dim=3, r=0.15)
    mf.initialize()
    return mf


def _horner(rowp, cols, B, D, N, x, transpose=False):
    """The partition-of-unity filter written out with scipy:
    T^-1 sum_{k<=N} (D^-1 B)^k D^-1 x, T^-1 = 1 / F(1)."""
    import scipy.sparse as sp
    n = len(D)
    Bm = sp.csr_matrix((B, cols, rowp), shape=(n, n))
    Dinv = 1.0 / D

    def fwd(v):
        t = Dinv * v
        y = t.copy()
        for _ in range(N):
            y = Dinv * (Bm @ y) + t
        return y
    Tinv = 1.0 / fwd(np.ones(n))
    if not transpose:
        return Tinv * fwd(x)
    t = Tinv * x
    y = t.copy()
    for _ in range(N):
        y = Bm.T @ (Dinv * y) + t
    return Dinv * y


def test_mfilter_matches_the_written_out_sum():
    """TopOptUtils.Mfilter on the GPU (tmrx_filter_create_custom): F x and
    F^T y equal the Horner sum written out with scipy from the same
    stencils; constants pass; the stencils meet their moment conditions."""
    forest = TMR.createBoxForest(2, 1, 1, L=(2.0, 1.0, 1.0), face_names=FACES)
    forest.createTrees(2)
    rng = np.random.default_rng(4)
    forest.refine(rng.integers(0, 2, len(forest.getOctants())).astype(np.intc), 0, 3)
    mat = TMR.MaterialProperties(rho=2.5, E=1.0, nu=0.3, ys=1.0)
    props = TMR.StiffnessProperties(mat, q=3.0, k0=1e-6)
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    def callback(f):
        c = topo.OctConformTopoCreator(bcs, f, props=props)
        return c, c.getFilter()

    prob = TopOptUtils.createTopoProblem(forest, callback, _mfilter, nlevels=2)
    mf = prob.getTopoFilter()
    rowp, cols, B, D = mf._stencil
    n = len(D)
    x, y = rng.random(n), rng.standard_normal(n)
    fx, fty = mf.applyFilter(x), mf.applyTranspose(y)
    rx, rty = _horner(rowp, cols, B, D, 4, x), _horner(rowp, cols, B, D, 4, y, True)
    e1 = np.linalg.norm(fx - rx) / np.linalg.norm(rx)
    e2 = np.linalg.norm(fty - rty) / np.linalg.norm(rty)
    print("Mfilter: %d nodes, %d stencil entries, rel err %.2e, transpose %.2e"
          % (n, len(cols), e1, e2), flush=True)
    assert e1 < 1e-12 and e2 < 1e-12
    assert np.allclose(mf.applyFilter(np.full(n, 0.7)), 0.7, rtol=1e-12, atol=0)
    assert abs(fx @ y - x @ fty) <= 1e-11 * abs(fx @ y)
    assert B.min() >= 0.0 and D.min() >= 1.0


@pytest.mark.parametrize("filter_type", ["helmholtz", "matrix", "mfilter"])
def test_topo_problem_with_filter(filter_type):
    """createTopoProblem(..., "helmholtz" | "matrix"): the analysis sees the
    filtered nodal design; the design-space gradient is the filter transpose
    of the nodal one and matches central differences."""
    forest = TMR.createBoxForest(2, 1, 1, L=(2.0, 1.0, 1.0), face_names=FACES)
    forest.createTrees(2)
    rng = np.random.default_rng(3)
    forest.refine(rng.integers(0, 2, len(forest.getOctants())).astype(np.intc), 0, 3)
    mat = TMR.MaterialProperties(rho=2.5, E=1.0, nu=0.3, ys=1.0)
    props = TMR.StiffnessProperties(mat, q=3.0, k0=1e-6)
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    def callback(f):
        c = topo.OctConformTopoCreator(bcs, f, props=props)
        return c, c.getFilter()

    ft = _mfilter if filter_type == "mfilter" else filter_type
    prob = TopOptUtils.createTopoProblem(forest, callback, ft, nlevels=2, r0=0.1, N=6)
    a = prob.getAssembler()
    f1 = TopOptUtils.compute3DTractionLoad("load", a.forest, a, [0.0, 0.0, -1.0])
    prob.setLoadCases([f1])
    vol = a.volumes.sum()
    prob.addConstraints(0, [topo.StructuralMass()], [-0.4 * vol * 2.5], [-1.0])
    prob.rtol = 1e-11
    x0, lb, ub = prob.getVarsAndBounds()
    x = np.clip(0.3 + 0.7 * rng.random(len(x0)), lb, ub)
    fail, fobj, cons = prob.evalObjCon(1, x)
    assert fail == 0
    g = np.zeros(a.ndv)
    A = [np.zeros(a.ndv)]
    prob.evalObjConGradient(x, g, A)
    d = rng.standard_normal(len(x))
    h = 1e-4
    fp, cp = prob.evalObjCon(1, x + h * d)[1:]
    fm, cm = prob.evalObjCon(1, x - h * d)[1:]
    fd, fdc = (fp - fm) / (2 * h), (cp[0] - cm[0]) / (2 * h)
    print("%s: objective fd %.10g g.d %.10g; mass fd %.10g A.d %.10g"
          % (filter_type, fd, g @ d, fdc, A[0] @ d), flush=True)
    assert abs(fd - g @ d) / abs(fd) < 1e-5
    assert abs(fdc - A[0] @ d) <= 1e-8 * abs(fdc)
