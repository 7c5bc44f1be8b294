"""The TMR-style optimisation problem on the GPU (TopOptUtils.createTopoProblem
-> TopoProblem.evalObjCon / evalObjConGradient) against the FP64 CPU oracle:
weighted multi-load-case compliance, mass and p-norm stress constraints and
their design-space gradients W^T dF/drho."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import TMR, TopOptUtils, topo

pytestmark = pytest.mark.gpu

FACES = {"x-": "fixed", "x+": "load", "z+": "top"}
Q, K0, P_NORM = 3.0, 1e-6, 8.0


def _problem(penalty="SIMP"):
    forest = TMR.createBoxForest(2, 1, 1, L=(2.0, 1.0, 1.0), face_names=FACES)
    forest.createTrees(2)
    rng = np.random.default_rng(11)
    forest.refine(rng.integers(0, 2, len(forest.getOctants())).astype(np.intc), 0, 4)
    mat = TMR.MaterialProperties(rho=2.5, E=1.0, nu=0.3, ys=1.0)
    props = TMR.StiffnessProperties(mat, q=Q, k0=K0, eps=0.3, penalty_type=penalty)
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    def callback(f):
        filt = f.coarsen()
        filt.balance(1)
        filt.createNodes()
        return topo.OctTopoCreator(bcs, filt, props=props), filt

    prob = TopOptUtils.createTopoProblem(forest, callback, "lagrange", nlevels=3)
    a = prob.getAssembler()
    f1 = TopOptUtils.compute3DTractionLoad("load", a.forest, a, [0.0, 0.0, -1.0])
    f2 = TopOptUtils.compute3DTractionLoad("top", a.forest, a, [0.0, 0.3, -0.5])
    prob.setLoadCases([f1, f2])
    return prob, a, props, [f1, f2]


def _oracle_case(a, rho, f, ptype):
    orc = po.from_forest(a.forest, [("fixed", 7)])
    orc.assemble(np.array([po.penalty(r, Q, K0, ptype) for r in rho]))
    u = orc.solve(f, rtol=1e-13)
    e = orc.energy(u)
    return orc, u, float(f @ u), -np.array([po.dpenalty(r, Q, ptype) for r in rho]) * e


@pytest.mark.parametrize("penalty", ["SIMP", "RAMP"])
def test_weighted_compliance_and_mass(penalty):
    prob, a, props, forces = _problem(penalty)
    ptype = props.penalty_type
    w = [1.0, 0.5]
    prob.setObjective(w)
    vol = a.volumes.sum()
    prob.addConstraints(0, [topo.StructuralMass()], [-0.4 * vol * 2.5], [-1.0])
    prob.rtol = 1e-11
    x0, lb, ub = prob.getVarsAndBounds()
    rng = np.random.default_rng(5)
    x = np.clip(0.3 + 0.7 * rng.random(len(x0)), lb, ub)
    print("topo", penalty, "nelem", a.nelem, "ndv", a.ndv, flush=True)
    fail, fobj, cons = prob.evalObjCon(1, x)
    assert fail == 0
    rho = a.element_densities(x)
    refs = [_oracle_case(a, rho, f, ptype) for f in forces]
    fref = sum(wi * r[2] for wi, r in zip(w, refs))
    print("fobj %.12g ref %.12g" % (fobj, fref), prob.last_info, flush=True)
    assert abs(fobj - fref) / abs(fref) < 1e-6
    mass = 2.5 * float(a.volumes @ rho)
    assert np.isclose(cons[0], -(mass - 0.4 * vol * 2.5), rtol=1e-12)
    g = np.zeros(a.ndv)
    A = [np.zeros(a.ndv)]
    assert prob.evalObjConGradient(x, g, A) == 0
    gref = a.weights_transpose(sum(wi * r[3] for wi, r in zip(w, refs)))
    err = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print("grad rel err %.3e" % err, flush=True)
    assert err < 1e-6
    assert np.allclose(A[0], -2.5 * a.weights_transpose(a.volumes), rtol=1e-12)
    # central difference of the objective along a random design direction
    d = rng.standard_normal(len(x))
    h = 1e-4
    fp = prob.evalObjCon(1, x + h * d)[1]
    fm = prob.evalObjCon(1, x - h * d)[1]
    fd = (fp - fm) / (2 * h)
    print("fd %.10g g.d %.10g" % (fd, g @ d), flush=True)
    assert abs(fd - g @ d) / abs(fd) < 1e-5


def test_stress_constraint_gradient():
    prob, a, props, forces = _problem("SIMP")
    prob.setObjective([1.0, 0.0])
    prob.addConstraints(1, [topo.PNormFailure(P_NORM)], [-2.0], [-0.5])
    prob.rtol = 1e-11
    x0, lb, ub = prob.getVarsAndBounds()
    x = np.clip(0.4 + 0.6 * np.random.default_rng(8).random(len(x0)), lb, ub)
    fail, fobj, cons = prob.evalObjCon(1, x)
    assert fail == 0
    rho = a.element_densities(x)
    orc = po.from_forest(a.forest, [("fixed", 7)])
    fs, gs, _, _ = po.stress_gradient(orc, rho, forces[1], q=Q, k0=K0, P=P_NORM, ys=1.0, eps=0.3)
    print("stress %.12g ref %.12g" % (-cons[0] / 0.5 + 2.0, fs), flush=True)
    assert abs(cons[0] - (-0.5) * (fs - 2.0)) / abs(cons[0]) < 1e-6
    g = np.zeros(a.ndv)
    A = [np.zeros(a.ndv)]
    prob.evalObjConGradient(x, g, A)
    Aref = -0.5 * a.weights_transpose(gs)
    err = np.linalg.norm(A[0] - Aref) / np.linalg.norm(Aref)
    print("stress grad rel err %.3e" % err, flush=True)
    assert err < 1e-6


def test_traction_load_matches_native():
    prob, a, props, forces = _problem("SIMP")
    fn = a.problem.traction_load("load", [0.0, 0.0, -1.0])
    mask = po.bcmask_from_forest(a.forest, [("fixed", 7)])
    fn.reshape(-1, 3)[mask == 7] = 0.0
    assert np.allclose(fn, forces[0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("penalty", ["SIMP", "RAMP"])
def test_conform_filter_nodal_design(penalty):
    """createTopoProblem(..., "conform") with OctConformTopoCreator levels:
    nodal design variables on the analysis forests, per-Gauss-point penalty,
    weighted compliance and mass with gradients per node."""
    forest = TMR.createBoxForest(2, 1, 1, L=(2.0, 1.0, 1.0), face_names=FACES)
    forest.createTrees(2)
    rng = np.random.default_rng(12)
    forest.refine(rng.integers(0, 2, len(forest.getOctants())).astype(np.intc), 0, 4)
    mat = TMR.MaterialProperties(rho=2.5, E=1.0, nu=0.3, ys=1.0)
    props = TMR.StiffnessProperties(mat, q=Q, k0=K0, eps=0.3, penalty_type=penalty)
    ptype = props.penalty_type
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    def callback(f):
        c = topo.OctConformTopoCreator(bcs, f, props=props)
        return c, c.getFilter()

    prob = TopOptUtils.createTopoProblem(forest, callback, "conform", nlevels=3)
    a = prob.getAssembler()
    assert a.nodal and a.ndv == a.forest.getNodeRange()[1]
    f1 = TopOptUtils.compute3DTractionLoad("load", a.forest, a, [0.0, 0.0, -1.0])
    f2 = TopOptUtils.compute3DTractionLoad("top", a.forest, a, [0.0, 0.3, -0.5])
    prob.setLoadCases([f1, f2])
    w = [1.0, 0.5]
    prob.setObjective(w)
    vol = a.volumes.sum()
    prob.addConstraints(0, [topo.StructuralMass()], [-0.4 * vol * 2.5], [-1.0])
    prob.rtol = 1e-11
    x0, lb, ub = prob.getVarsAndBounds()
    x = np.clip(0.3 + 0.7 * rng.random(len(x0)), lb, ub)
    print("conform", penalty, "nelem", a.nelem, "ndv", a.ndv, flush=True)
    fail, fobj, cons = prob.evalObjCon(1, x)
    assert fail == 0
    orc = po.from_forest(a.forest, [("fixed", 7)])
    refs = [po.nodal_compliance_gradient(orc, x, f, Q, K0, ptype) for f in (f1, f2)]
    fref = sum(wi * r[0] for wi, r in zip(w, refs))
    print("fobj %.12g ref %.12g" % (fobj, fref), prob.last_info, flush=True)
    assert abs(fobj - fref) / abs(fref) < 1e-6
    # mass: exact integral of the trilinear density
    rg = orc.gp_density(x)
    mass = 2.5 * float(np.sum(a.volumes[:, None] / 8.0 * rg))
    assert np.isclose(cons[0], -(mass - 0.4 * vol * 2.5), rtol=1e-12)
    g = np.zeros(a.ndv)
    A = [np.zeros(a.ndv)]
    assert prob.evalObjConGradient(x, g, A) == 0
    gref = sum(wi * r[1] for wi, r in zip(w, refs))
    err = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print("grad rel err %.3e" % err, flush=True)
    assert err < 1e-6
    assert np.allclose(A[0], -2.5 * a.mass_weights(), rtol=1e-12)
    d = rng.standard_normal(len(x))
    h = 1e-4
    fp = prob.evalObjCon(1, x + h * d)[1]
    fm = prob.evalObjCon(1, x - h * d)[1]
    fd = (fp - fm) / (2 * h)
    print("fd %.10g g.d %.10g" % (fd, g @ d), flush=True)
    assert abs(fd - g @ d) / abs(fd) < 1e-5
