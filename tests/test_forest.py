"""The native forest (tmr_amd/csrc/forest.cpp, host code, no GPU) against
the geometric oracle: octants after balance, element connectivity, hanging
node maps and weights, interpolation CSR, named nodes -- bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import TMR, _lib

KEY = ["block", "x", "y", "z", "level"]


def build(dims, level, seed=None, maxlev=5, corner=1, frac=1.0):
    L = (float(dims[0]), float(dims[1]), float(dims[2]))
    f = TMR.createBoxForest(*dims, L=L, face_names={"x-": "fixed", "x+": "load", "z+": "top"})
    f.createTrees(level)
    if seed is not None:
        rng = np.random.default_rng(seed)
        n = len(f.getOctants())
        r = (rng.integers(0, 3, n) * (rng.random(n) < frac)).astype(np.intc)
        f.refine(r, 0, maxlev)
    pre = f.getOctants()
    f.balance(corner)
    f.createNodes()
    return f, pre


@pytest.mark.parametrize("dims,level", [((1, 1, 1), 3), ((2, 1, 1), 3), ((2, 2, 2), 2), ((3, 1, 2), 2)])
def test_uniform_matches_oracle(dims, level):
    f, pre = build(dims, level)
    m = po.OracleMesh(dims, pre).create_nodes()
    assert np.array_equal(m.leaves[KEY], f.getOctants()[KEY])
    assert np.array_equal(m.conn, f.getMeshConn())
    assert f.getNodeRange()[1] == m.nindep
    assert np.all(f.getOctants()["tag"] == np.arange(len(m.leaves)))


@pytest.mark.parametrize("seed,corner", [(0, 1), (1, 1), (2, 1), (3, 0), (4, 0), (5, 1)])
def test_adaptive_maps_match_oracle(seed, corner):
    dims = (2, 2, 2)
    f, pre = build(dims, 1, seed=seed, maxlev=5, corner=corner)
    m = po.OracleMesh(dims, pre, corner).create_nodes()
    assert np.array_equal(m.leaves[KEY], f.getOctants()[KEY])
    assert np.array_equal(m.conn, f.getMeshConn())
    ptr, dc, dw = f.getDepNodeConn()
    assert m.ndep == len(ptr) - 1 and m.ndep > 0
    assert np.array_equal(m.dep_ptr, ptr)
    assert np.array_equal(m.dep_conn, dc)
    assert np.array_equal(m.dep_w, dw)


@pytest.mark.parametrize("seed", [0, 7])
def test_interpolation_matches_oracle(seed):
    dims = (2, 2, 2)
    f, pre = build(dims, 1, seed=seed, maxlev=5)
    c = f.coarsen()
    cpre = c.getOctants()
    c.balance(1)
    c.createNodes()
    rp, cl, w = f.createInterpolation(c)
    fm, cm, (orp, ocl, ow) = po.interpolation(dims, pre, cpre)
    assert np.array_equal(rp, orp)
    assert np.array_equal(cl, ocl)
    assert np.array_equal(w, ow)


def test_named_nodes_match_oracle():
    dims = (2, 2, 2)
    f, pre = build(dims, 1, seed=3, maxlev=4)
    m = po.OracleMesh(dims, pre).create_nodes()
    assert np.array_equal(f.getNodesWithName("fixed"), m.nodes_on_box_face(0, 0))
    assert np.array_equal(f.getNodesWithName("load"), m.nodes_on_box_face(0, 1))
    assert np.array_equal(f.getNodesWithName("top"), m.nodes_on_box_face(2, 1))
    # hanging nodes on a named face are reported with their negative numbers
    # (TMROctForest::getNodesWithName copies conn entries, :6098-6167)
    assert np.any(f.getNodesWithName("fixed") < 0) == np.any(m.nodes_on_box_face(0, 0) < 0)


def test_octs_with_name_face_index_and_tag():
    f, _ = build((2, 1, 1), 2)
    octs = f.getOctsWithName("load")
    assert len(octs) == 4 * 4  # one block face at level 2
    assert np.all(octs["info"] == 1)
    allo = f.getOctants()
    assert np.all(allo["x"][octs["tag"]] + (1 << 28) == (1 << 30))
    assert np.all(allo["block"][octs["tag"]] == 1)


def test_points_are_lattice_coordinates():
    f, _ = build((2, 1, 1), 2)
    X = f.getPoints()
    nums = f.getNodeNumbers()
    assert len(X) == len(nums)
    conn = f.getMeshConn()
    ext = f.getElementExtents()
    assert np.allclose(ext, 0.25)
    # every element's corner 7 minus corner 0 is its diagonal
    nd = len(f.getDepNodeConn()[0]) - 1
    idx = lambda n: n + nd if n >= 0 else nd - 1 - (-n - 1)
    for e in range(0, len(conn), 7):
        d = X[idx(conn[e, 7])] - X[idx(conn[e, 0])]
        assert np.allclose(d, ext[e])


def test_coarsen_takes_parents_of_zero_children():
    f, _ = build((1, 1, 1), 3)
    c = f.coarsen()
    o = c.getOctants()
    assert len(o) == 64 and np.all(o["level"] == 2)


def test_errors_are_reported():
    f = TMR.OctForest()
    with pytest.raises(_lib.TmrxError):
        f.createTrees(2)  # no connectivity
    g = TMR.createBoxForest(1, 1, 1)
    g.createTrees(1)
    with pytest.raises(ValueError):
        g.refine(np.zeros(3, np.intc))
    # an unbalanced forest cannot number its nodes
    r = np.zeros(8, np.intc)
    r[0] = 3
    g.refine(r, 0, 6)
    with pytest.raises(_lib.TmrxError):
        g.createNodes()
    with pytest.raises(ValueError):
        g.createInterpolation(g)


@pytest.mark.parametrize("seed", [0, 4])
def test_element_center_weights_reproduce_linear_fields(seed):
    """TMROctTACSTopoCreator's element densities (src/topology/
    TMR_TACSTopoCreator.cpp:55-186) are the trilinear filter field at each
    element centre: rows form a partition of unity and reproduce any linear
    field sampled at the filter's independent nodes (hanging-node weights
    are exact for linear fields)."""
    dims = (2, 2, 2)
    f, _ = build(dims, 1, seed=seed, maxlev=5)
    c = f.coarsen()
    c.balance(1)
    c.createNodes()
    rp, cl, w, nc = f.createElementCenterWeights(c)
    assert len(rp) - 1 == len(f.getOctants()) and nc == c.getNodeRange()[1]
    assert np.all(np.diff(rp) >= 1) and np.all((cl >= 0) & (cl < nc))
    for e in range(len(rp) - 1):  # sorted, unique columns per row
        assert np.all(np.diff(cl[rp[e]:rp[e + 1]]) > 0)
    assert np.allclose(np.add.reduceat(w, rp[:-1]), 1.0, rtol=0, atol=1e-14)
    ndc = len(c.getDepNodeConn()[0]) - 1
    Xc = c.getPoints()[ndc:]
    ndf = len(f.getDepNodeConn()[0]) - 1
    Xf = f.getPoints()
    conn = f.getMeshConn()
    idx = np.where(conn >= 0, conn + ndf, ndf - 1 - (-conn - 1))
    centres = Xf[idx].mean(axis=1)
    a = np.array([0.3, -1.7, 2.2])
    lin = Xc @ a + 0.5
    rho = np.zeros(len(rp) - 1)
    np.add.at(rho, np.repeat(np.arange(len(rp) - 1), np.diff(rp)), w * lin[cl])
    assert np.allclose(rho, centres @ a + 0.5, rtol=0, atol=1e-12)


def test_element_center_weights_same_forest():
    """With the analysis forest as its own filter, each element centre takes
    1/8 of each corner (hanging corners expanded through their parents)."""
    f, _ = build((1, 1, 1), 2)
    g = f.duplicate()
    g.createNodes()
    rp, cl, w, nc = f.createElementCenterWeights(g)
    assert np.all(np.diff(rp) == 8) and np.all(w == 0.125)
    assert np.array_equal(cl.reshape(-1, 8), np.sort(f.getMeshConn(), axis=1))


def test_c2_forest_maps_bit_exact():
    """The adaptive benchmark forest at full size (1,227,458 hex, levels 3-7,
    2:1 balanced with corners) and its first coarse level: element
    connectivity, hanging-node pointers, parents and weights, and the
    interpolation CSR equal the geometric oracle's bit for bit."""
    from tmr_amd import workloads
    forests, rho, L = workloads.c2_adaptive(seed=0, nlevels=2)
    fine, coarse = forests
    assert len(fine.getOctants()) == 1227458
    fm, cm, (rp, cl, w) = po.interpolation((2, 2, 2), fine.getOctants(), coarse.getOctants())
    assert np.array_equal(np.asarray(fine.getMeshConn()).reshape(-1, 8), fm.conn)
    dp, dc, dw = fine.getDepNodeConn()
    assert len(dp) - 1 == fm.ndep and fm.ndep > 100000
    assert np.array_equal(dp, fm.dep_ptr)
    assert np.array_equal(dc, fm.dep_conn)
    assert np.array_equal(dw, fm.dep_w)
    prp, pcl, pw = fine.createInterpolation(coarse)
    assert np.array_equal(prp, rp)
    assert np.array_equal(pcl, cl)
    assert np.array_equal(pw, w)


def test_set_mesh_order_and_local_node_count():
    """setMeshOrder frees the nodes (they are rebuilt by createNodes) and
    only order 2 is built; getNodeConn's local-node count is the length of
    the node numbers (owned independent plus dependent nodes here)."""
    import ctypes as C
    f, pre = build((2, 2, 2), 1, seed=3, maxlev=4)
    conn0 = f.getMeshConn().copy()
    f.setMeshOrder(2)
    with pytest.raises(_lib.TmrxError):
        f.getMeshConn()
    f.createNodes()
    assert np.array_equal(f.getMeshConn(), conn0)
    with pytest.raises(_lib.TmrxError, match="order 2"):
        f.setMeshOrder(3)
    f.setMeshOrder(1)  # clamped to 2, as the reference does
    f.createNodes()
    p = C.POINTER(C.c_int)()
    ne, no, nl = C.c_int(), C.c_int(), C.c_int()
    assert _lib.lib().tmrx_forest_get_node_conn(f.ptr, C.byref(p), C.byref(ne), C.byref(no),
                                                C.byref(nl)) == 0
    dp, dc, dw = f.getDepNodeConn()
    assert ne.value == len(f.getOctants())
    assert no.value == f.getNodeRange()[1]
    assert nl.value == len(f.getNodeNumbers()) == no.value + len(dp) - 1


# Block connectivities with rotated faces (examples/parallel/octant_test.cpp:
# the two-block rectangle, whose second block has its own axes, and the
# seven-block box, a cube inside a cube).  The geometric oracle covers
# equally oriented lattices only; these fixtures are pinned by closed forms.
RECT_X = np.array([0, 0, 0, 1, 0, 0, 2, 0, 0, 0, 1, 0, 1, 1, 0, 2, 1, 0,
                   0, 0, 1, 1, 0, 1, 2, 0, 1, 0, 1, 1, 1, 1, 1, 2, 1, 1], float).reshape(-1, 3)
RECT_CONN = np.array([[0, 1, 3, 4, 6, 7, 9, 10], [8, 11, 2, 5, 7, 10, 1, 4]], np.intc)
BOX_X = np.array([-.5, -.5, -.5, .5, -.5, -.5, -.5, .5, -.5, .5, .5, -.5, -.5, -.5, .5, .5,
                  -.5, .5, -.5, .5, .5, .5, .5, .5, -1, -1, -1, -1, -1, 1, 1, -1,
                  -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, 1, 1, -1, 1, 1, 1]).reshape(-1, 3)
BOX_CONN = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 0, 1, 9, 11, 4, 5, 5, 11, 1, 10, 7, 15, 3,
                     14, 7, 15, 3, 14, 6, 13, 2, 12, 9, 13, 4, 6, 8, 12, 0, 2, 10, 14, 8, 12,
                     1, 3, 0, 2, 4, 5, 6, 7, 9, 11, 13, 15], np.intc).reshape(-1, 8)


def _rotated_forest(X, conn, level, refine_seed=None):
    f = TMR.OctForest()
    f.setConnectivity(conn)
    f.setNodePoints(X)
    f.createTrees(level)
    if refine_seed is not None:
        rng = np.random.default_rng(refine_seed)
        f.refine((rng.random(len(f.getOctants())) < 0.2).astype(np.intc), 0, level + 2)
    f.balance(1)
    f.createNodes()
    return f


def _indep_points(f):
    """Coordinates of the independent nodes 0..n-1 (getPoints follows the
    sorted node numbers, dependent nodes first)."""
    nums = f.getNodeNumbers()
    X = f.getPoints()
    n = int(f.getNodeRange()[1])
    out = np.zeros((n, 3))
    sel = nums >= 0
    out[nums[sel]] = X[sel]
    return out, X[~sel], nums[~sel]


@pytest.mark.parametrize("which", ["rectangle", "box"])
@pytest.mark.parametrize("level", [1, 2])
def test_rotated_blocks_uniform_node_count(which, level):
    """Conforming uniform mesh: V + E m + F m^2 + C m^3 nodes, m = 2^L - 1,
    with V, E, F, C the block vertices, edges, faces and blocks (rectangle
    12, 20, 11, 2; box 16, 32, 24, 7: Euler's V - E + F - C = 1 and
    F = (6 C + 6) / 2 for a cube-shaped complex); no hanging nodes."""
    X, conn = (RECT_X, RECT_CONN) if which == "rectangle" else (BOX_X, BOX_CONN)
    f = _rotated_forest(X, conn, level)
    V, E, F, Cb = (12, 20, 11, 2) if which == "rectangle" else (16, 32, 24, 7)
    assert V - E + F - Cb == 1
    m = 2 ** level - 1
    assert f.getNodeRange()[1] == V + E * m + F * m * m + Cb * m ** 3
    assert len(f.getDepNodeConn()[0]) == 1


@pytest.mark.parametrize("which", ["rectangle", "box"])
def test_rotated_blocks_hanging_nodes_and_interpolation(which):
    """Adaptive refinement across rotated faces: every hanging node sits at
    the weighted mean of its parents (weights 1/2 or 1/4 summing to 1), and
    the prolongation from the coarsened forest reproduces an affine field of
    the physical coordinates exactly (trilinear block maps keep affine fields
    trilinear in every element); on the rectangle, whose blocks are boxes, a
    global trilinear field with cross terms too."""
    X, conn = (RECT_X, RECT_CONN) if which == "rectangle" else (BOX_X, BOX_CONN)
    f = _rotated_forest(X, conn, 2, refine_seed=4)
    P_ind, P_dep, dep_nums = _indep_points(f)
    dp, dc, dw = f.getDepNodeConn()
    assert len(dp) - 1 > 0
    for k, d in enumerate(-dep_nums - 1):
        w = dw[dp[d]:dp[d + 1]]
        assert np.all((w == 0.5) | (w == 0.25)) and abs(w.sum() - 1.0) < 1e-15
        mean = (w[:, None] * P_ind[dc[dp[d]:dp[d + 1]]]).sum(0)
        assert np.allclose(P_dep[k], mean, atol=1e-14)
    c = f.coarsen()
    c.balance(1)
    c.createNodes()
    rp, cl, w = f.createInterpolation(c)
    Pc, _, _ = _indep_points(c)
    fields = [lambda p: 0.3 - 1.25 * p[:, 0] + 2.0 * p[:, 1] + 0.5 * p[:, 2]]
    if which == "rectangle":
        fields.append(lambda p: (0.3 - p[:, 0] + 0.7 * p[:, 1] * p[:, 2] +
                                 1.1 * p[:, 0] * p[:, 1] - 0.6 * p[:, 0] * p[:, 1] * p[:, 2]))
    for fld in fields:
        Pu = np.add.reduceat(w * fld(Pc)[cl], rp[:-1])
        assert np.max(np.abs(Pu - fld(P_ind))) < 1e-13
