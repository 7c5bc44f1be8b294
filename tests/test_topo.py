"""Host logic of the TMR-style topology layer (tmr_amd/topo.py,
tmr_amd/TopOptUtils.py): boundary conditions, stiffness parametrisation,
creator weights and load vectors.  No GPU."""
import numpy as np
import pytest

from tmr_amd import TMR, TopOptUtils, topo

FACES = {"x-": "fixed", "x+": "load", "z+": "top"}


def _forest(level=3, seed=None, dims=(2, 1, 1)):
    f = TMR.createBoxForest(*dims, L=tuple(float(d) for d in dims), face_names=FACES)
    f.createTrees(level)
    if seed is not None:
        rng = np.random.default_rng(seed)
        f.refine(rng.integers(0, 2, len(f.getOctants())).astype(np.intc), 0, level + 2)
    f.balance(1)
    f.createNodes()
    return f


def _creator(forest, props, bcs):
    filt = forest.coarsen()
    filt.balance(1)
    filt.createNodes()
    return topo.OctTopoCreator(bcs, filt, props=props), filt


def test_boundary_conditions():
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")
    bcs.addBoundaryCondition("sym", [0, 2])
    bcs.addBoundaryCondition("roller", [1], [0.0])
    assert bcs.getNumBoundaryConditions() == 3
    assert bcs.as_list() == [("fixed", 7), ("sym", 5), ("roller", 2)]
    with pytest.raises(ValueError):
        bcs.addBoundaryCondition("x", [0, 1], [0.0])
    with pytest.raises(ValueError):
        bcs.addBoundaryCondition("x", [3])
    assert not bcs.has_values()
    # prescribed values: one per listed component (tmr/TMR.pyx:3917-3927)
    bcs.addBoundaryCondition("load", [2, 0], [0.25, -1.5])
    assert bcs.has_values()
    assert bcs.as_list()[-1] == ("load", 5)
    forest = _forest()
    u = bcs.values_vector(forest)
    nodes = np.asarray(forest.getNodesWithName("load"))
    nodes = nodes[nodes >= 0]
    assert np.all(u[3 * nodes] == -1.5) and np.all(u[3 * nodes + 2] == 0.25)
    assert np.all(u[3 * nodes + 1] == 0.0)
    assert np.count_nonzero(u) == 2 * len(nodes)


def test_stiffness_properties_defaults_and_aliases():
    m = TMR.MaterialProperties(rho=2.0, E=70e3, nu=0.3, ys=100.0)
    p = TMR.StiffnessProperties(m)
    # tmr/TMR.pyx:5047-5057 defaults
    assert (p.q, p.eps, p.k0, p.penalty_type) == (5.0, 0.3, 1e-6, TMR.RAMP_PENALTY)
    assert p.design_lower_bound() == 0.0
    s = TMR.StiffnessProperties([m], stiffness_penalty_value=3.0, stress_relax_value=0.1,
                                stiffness_offset=1e-4, penalty_type="SIMP")
    assert (s.q, s.eps, s.k0, s.penalty_type) == (3.0, 0.1, 1e-4, TMR.SIMP_PENALTY)
    assert s.design_lower_bound() == 1e-3
    with pytest.raises(NotImplementedError):
        TMR.StiffnessProperties([m, m])


def test_creator_assembler_weights_and_callback():
    forest = _forest(seed=1)
    props = TMR.StiffnessProperties(TMR.MaterialProperties())
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    class Creator(TMR.OctTopoCreator):
        calls = 0

        def createElement(self, order, octant, index, weights):
            Creator.calls += 1
            assert order == 2 and len(index) == len(weights)
            return TMR.OctConstitutive(props, self.getFilter())

    filt = forest.coarsen()
    filt.balance(1)
    filt.createNodes()
    a = Creator(bcs, filt).createTACS(forest)
    assert Creator.calls == 1 and a.props is props
    assert a.nelem == len(forest.getOctants()) and a.ndv == filt.getNodeRange()[1]
    x = a.getDesignVars()
    assert np.all(x == 0.95)
    assert np.allclose(a.element_densities(x), 0.95, rtol=0, atol=1e-14)
    lb, ub = a.getDesignVarRange()
    assert np.all(lb == 0.0) and np.all(ub == 1.0)
    # W^T is the adjoint of W
    rng = np.random.default_rng(0)
    xr, gr = rng.random(a.ndv), rng.random(a.nelem)
    assert np.isclose(a.element_densities(xr) @ gr, xr @ a.weights_transpose(gr), rtol=1e-13)
    assert np.isclose(a.volumes.sum(), 2.0)
    with pytest.raises(ValueError):
        a.element_densities(np.zeros(a.ndv + 1))


def test_conform_creator_nodal_assembler():
    """OctConformTopoCreator: element i's design variables are its nodes
    (TMR_TACSTopoCreator.cpp:714-777); mass weights integrate the trilinear
    density exactly."""
    forest = _forest(seed=2)
    props = TMR.StiffnessProperties(TMR.MaterialProperties(rho=2.0), penalty_type="SIMP")
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")

    class Creator(TMR.OctConformTopoCreator):
        calls = 0

        def createElement(self, order, octant, index, filtr):
            Creator.calls += 1
            assert order == 2 and len(index) == 8 and filtr is self.getFilter()
            return TMR.OctConstitutive(props, filtr)

    a = Creator(bcs, forest).createTACS(forest)
    assert Creator.calls == 1 and a.nodal and a.props is props
    assert a.ndv == forest.getNodeRange()[1]
    lb, ub = a.getDesignVarRange()
    assert np.all(lb == 1e-3) and np.all(ub == 1.0)
    # a constant density c integrates to c * volume; a linear field exactly
    w = a.mass_weights()
    assert np.isclose(w.sum(), a.volumes.sum(), rtol=1e-14)
    nums = np.asarray(forest.getNodeNumbers())
    X = forest.getPoints()[nums >= 0]   # points in sorted node-number order
    assert np.array_equal(nums[nums >= 0], np.arange(a.ndv))
    lin = 0.5 + 0.25 * X[:, 0] - 0.125 * X[:, 2]
    # exact integral of 0.5 + 0.25 x - 0.125 z over [0,2]x[0,1]x[0,1]
    assert np.isclose(w @ lin, 0.5 * 2 + 0.25 * 2.0 - 0.125 * 1.0, rtol=1e-13)
    filt = TMR.ConformFilter([a], [forest])
    x = np.linspace(0.2, 0.9, a.ndv)
    assert np.array_equal(filt.applyFilter(x), x)
    assert np.array_equal(filt.addValues(x), x)
    with pytest.raises(ValueError):
        a.element_densities(x)
    with pytest.raises(NotImplementedError):
        TMR.OctConformTopoCreator(bcs, forest, design_vars_per_node=2)
    with pytest.raises(ValueError):
        TMR.ConformFilter([a], [])


@pytest.mark.parametrize("seed", [None, 2])
def test_traction_load_closed_form(seed):
    """Total force of a constant traction = face area x traction; on a
    uniform face the consistent load is t h^2 at interior nodes, half on
    edges, a quarter at corners; clamped rows are zero."""
    forest = _forest(level=2, seed=seed)
    bcs = TMR.BoundaryConditions()
    bcs.addBoundaryCondition("fixed")
    a, _ = _creator(forest, TMR.StiffnessProperties(TMR.MaterialProperties()), bcs)
    a = a.createTACS(forest)
    t = np.array([0.5, -1.0, 2.0])
    f = TopOptUtils.compute3DTractionLoad("load", forest, a, t)
    assert np.allclose(f.reshape(-1, 3).sum(axis=0), 1.0 * t, rtol=1e-13)
    top = TopOptUtils.compute3DTractionLoad("top", forest, a, t)
    fixed = forest.getNodesWithName("fixed")
    fixed = fixed[fixed >= 0]
    assert np.all(top.reshape(-1, 3)[fixed] == 0.0)
    # the clamped x=0 edge of the top face takes its share away from the total
    if seed is None:
        h = 0.25
        nodes = forest.getNodesWithName("load")
        X = forest.getPoints()[len(forest.getDepNodeConn()[0]) - 1:]
        on_edge = lambda n: (np.isclose(X[n, 1], 0) | np.isclose(X[n, 1], 1)).astype(int) + \
            (np.isclose(X[n, 2], 0) | np.isclose(X[n, 2], 1)).astype(int)
        share = np.array([1.0, 0.5, 0.25])[on_edge(nodes)]
        assert np.allclose(f.reshape(-1, 3)[nodes], share[:, None] * h * h * t, rtol=1e-13)
    # per-face tractions: only the x+ face (index 1) carries the load
    trac = [np.zeros(3)] * 6
    trac[1] = t
    g = TopOptUtils.computeTractionLoad(["load"], forest, a, trac)
    assert np.array_equal(g, f)
    with pytest.raises(ValueError):
        TopOptUtils.computeTractionLoad("load", forest, a, [t] * 5)


def test_vertex_load():
    forest = _forest(level=2)
    bcs = TMR.BoundaryConditions()
    a = topo.OctTopoCreator(bcs, forest.duplicate(), props=TMR.StiffnessProperties(
        TMR.MaterialProperties())).createTACS(forest)
    f = TopOptUtils.computeVertexLoad("load", forest, a, [0.0, 0.0, -1.0])
    nodes = forest.getNodesWithName("load")
    assert np.isclose(f.sum(), -float(np.sum(nodes >= 0)))
    with pytest.raises(ValueError):
        TopOptUtils.computeVertexLoad("load", forest, a, [0.0, 1.0])


def test_density_based_refine_rule():
    """TopOptUtils.densityBasedRefine (tmr/TopOptUtils.py:397-457): refine at
    or above upper, coarsen at or below lower (max of an element's values;
    reversed: min, and the roles swap)."""
    f = TMR.createBoxForest(1, 1, 1, face_names={"x-": "fixed"})
    f.createTrees(1)
    d = np.array([0.0, 0.05, 0.2, 0.5, 0.9, 0.04, 0.51, 0.3])
    flags = TopOptUtils.densityBasedRefine(f, d, lower=0.05, upper=0.5)
    assert flags.tolist() == [-1, -1, 0, 1, 1, -1, 1, 0]
    f2 = TMR.createBoxForest(1, 1, 1, face_names={"x-": "fixed"})
    f2.createTrees(1)
    dd = np.stack([d, np.ones_like(d)], 1)
    flags2 = TopOptUtils.densityBasedRefine(f2, dd, lower=0.05, upper=0.5, reverse=True)
    assert flags2.tolist() == [1, 1, 0, -1, -1, 1, -1, 0]
    assert f.getOctants()["level"].max() == 2
