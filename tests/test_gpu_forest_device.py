"""Forest passes on the GPU against the host passes (which test_forest.py
holds bit-exact against the geometric oracle):

* balance (tmrx_forest_balance_device): octants, order and tags;
* createNodes (tmrx_forest_create_nodes_device): octant info bits, element
  connectivity, hanging-node constraints, node numbers and points;
* createInterpolation (tmrx_forest_create_interpolation_device): the CSR;

all bit for bit, on uniform, adaptive (face- and corner-balanced),
eight-block, rotated-block and coarsened forests, and on the 1.2M-hex
adaptive hierarchy (built end to end on the GPU) and the 4.2M-hex box, where
the host and device passes are timed side by side."""
import time

import numpy as np
import pytest

from tmr_amd import TMR, workloads

pytestmark = pytest.mark.gpu

FACES = {"x-": "fixed", "x+": "load", "z+": "top"}


def _compare(f_host, f_dev, what):
    """f_host has host nodes; f_dev (same octants) gets device nodes."""
    t = time.time()
    ms = f_dev.createNodes(device=0)
    wall = 1e3 * (time.time() - t)
    a, b = f_host.getOctants(), f_dev.getOctants()
    assert a.dtype == b.dtype and np.array_equal(a, b), what
    assert np.array_equal(f_host.getMeshConn(), f_dev.getMeshConn()), what
    for x, y in zip(f_host.getDepNodeConn(), f_dev.getDepNodeConn()):
        assert np.array_equal(x, y), what
    assert np.array_equal(f_host.getNodeNumbers(), f_dev.getNodeNumbers()), what
    assert np.array_equal(f_host.getNodeRange(), f_dev.getNodeRange()), what
    Xa, Xb = f_host.getPoints(), f_dev.getPoints()
    assert Xa.shape == Xb.shape and np.array_equal(Xa, Xb), what
    print("%s: %d hex, %d nodes, %d hanging: identical (device pass %.1f ms, call %.1f ms)"
          % (what, len(a), len(f_dev.getNodeNumbers()), len(f_dev.getDepNodeConn()[0]) - 1,
             ms, wall), flush=True)


def _pair(f):
    """The forest (balanced, no nodes yet) and a duplicate of it."""
    g = f.duplicate()
    return f, g


def _box(dims, level, seed=None, maxlev=5, corner=1):
    f = TMR.createBoxForest(*dims, L=tuple(float(d) for d in dims), face_names=FACES)
    f.createTrees(level)
    if seed is not None:
        rng = np.random.default_rng(seed)
        f.refine(rng.integers(0, 3, len(f.getOctants())).astype(np.intc), 0, maxlev)
    f.balance(corner)
    return f


@pytest.mark.parametrize("dims,level", [((1, 1, 1), 3), ((2, 1, 1), 4), ((3, 1, 2), 2)])
def test_uniform(dims, level):
    f, g = _pair(_box(dims, level))
    f.createNodes()
    _compare(f, g, "uniform %s level %d" % (dims, level))


@pytest.mark.parametrize("seed,corner", [(1, 1), (2, 0), (5, 1)])
def test_adaptive(seed, corner):
    f, g = _pair(_box((2, 2, 1), 2, seed=seed, corner=corner))
    f.createNodes()
    _compare(f, g, "adaptive seed %d balance(%d)" % (seed, corner))


def test_eight_blocks_and_coarse_levels():
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=5, nlevels=3)
    for l, f in enumerate(forests):
        g = f.duplicate()
        _compare(f, g, "eight-block adaptive, level %d of the hierarchy" % l)


@pytest.mark.parametrize("which", ["rectangle", "box"])
def test_rotated_blocks(which):
    from test_forest import BOX_CONN, BOX_X, RECT_CONN, RECT_X
    X, conn = (RECT_X, RECT_CONN) if which == "rectangle" else (BOX_X, BOX_CONN)
    f = TMR.OctForest()
    f.setConnectivity(conn)
    f.setNodePoints(X)
    f.createTrees(2)
    rng = np.random.default_rng(4)
    f.refine((rng.random(len(f.getOctants())) < 0.2).astype(np.intc), 0, 4)
    f.balance(1)
    g = f.duplicate()
    f.createNodes()
    _compare(f, g, "rotated %s" % which)
    # the device-built forest serves the interpolation like the host one
    c = f.coarsen()
    c.balance(1)
    c.createNodes()
    for x, y in zip(f.createInterpolation(c), g.createInterpolation(c)):
        assert np.array_equal(x, y)


def test_large_forests_and_timing():
    t = time.time()
    forests, rho, L = workloads.c2_adaptive(seed=0, nlevels=1)
    f = forests[0]
    g = f.duplicate()
    print("C2 forest built in %.1f s (host nodes included)" % (time.time() - t), flush=True)
    _compare(f, g, "adaptive 1.2M-hex forest")
    b = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names=workloads.FACES)
    b.createTrees(7)
    b.balance(1)
    g = b.duplicate()
    t = time.time()
    b.createNodes()
    print("box 256x128x128: host createNodes %.0f ms" % (1e3 * (time.time() - t)), flush=True)
    _compare(b, g, "box 4.2M-hex forest")


def _unbalanced(dims, level, seed, maxlev=5):
    f = TMR.createBoxForest(*dims, L=tuple(float(d) for d in dims), face_names=FACES)
    f.createTrees(level)
    rng = np.random.default_rng(seed)
    f.refine(rng.integers(0, 3, len(f.getOctants())).astype(np.intc), 0, maxlev)
    return f


@pytest.mark.parametrize("seed,corner", [(1, 1), (2, 0), (7, 1)])
def test_balance(seed, corner):
    f = _unbalanced((2, 2, 1), 2, seed)
    g = f.duplicate()
    f.balance(corner)
    ms = g.balance(corner, device=0)
    a, b = f.getOctants(), g.getOctants()
    print("balance(%d) seed %d: %d octants, identical: %s (device %.1f ms)"
          % (corner, seed, len(a), np.array_equal(a, b), ms), flush=True)
    assert np.array_equal(a, b)
    # and the coarsened forest, which overlaps until balanced
    c = f.coarsen()
    d = c.duplicate()
    c.balance(1)
    d.balance(1, device=0)
    assert np.array_equal(c.getOctants(), d.getOctants())


@pytest.mark.parametrize("which", ["rectangle", "box"])
def test_balance_rotated_blocks(which):
    from test_forest import BOX_CONN, BOX_X, RECT_CONN, RECT_X
    X, conn = (RECT_X, RECT_CONN) if which == "rectangle" else (BOX_X, BOX_CONN)
    f = TMR.OctForest()
    f.setConnectivity(conn)
    f.setNodePoints(X)
    f.createTrees(2)
    rng = np.random.default_rng(9)
    f.refine(rng.integers(0, 3, len(f.getOctants())).astype(np.intc), 0, 5)
    g = f.duplicate()
    f.balance(1)
    g.balance(1, device=0)
    assert np.array_equal(f.getOctants(), g.getOctants())


def test_interpolation():
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=5, nlevels=3)
    for l in range(len(forests) - 1):
        a = forests[l].createInterpolation(forests[l + 1])
        b = forests[l].createInterpolation(forests[l + 1], device=0)
        for x, y in zip(a, b):
            assert x.dtype == y.dtype and np.array_equal(x, y)
    # between forests that are not one coarsening apart (design transfer):
    # the refined forest against the original
    orig = forests[0]
    new = orig.duplicate()
    rng = np.random.default_rng(2)
    new.refine(rng.integers(-1, 2, len(new.getOctants())).astype(np.intc), 0, 6)
    new.balance(1)
    new.createNodes()
    a = new.createInterpolation(orig)
    b = new.createInterpolation(orig, device=0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_c2_hierarchy_end_to_end_on_the_gpu():
    """The adaptive 1.2M-hex hierarchy (refinement sweeps, balance, coarsen,
    nodes, interpolation) with every balance, node and interpolation pass on
    the GPU, against the same hierarchy built on the host."""
    t = time.time()
    fh, rho_h, L = workloads.c2_adaptive(seed=0, nlevels=5)
    t_host = time.time() - t
    ph = [fh[l].createInterpolation(fh[l + 1]) for l in range(4)]
    t_host_all = time.time() - t
    t = time.time()
    fd, rho_d, L = workloads.c2_adaptive(seed=0, nlevels=5, device=0)
    t_dev = time.time() - t
    pd = [fd[l].createInterpolation(fd[l + 1], device=0) for l in range(4)]
    t_dev_all = time.time() - t
    print("C2 hierarchy: host %.2f s (+ interpolation %.2f s), GPU passes %.2f s "
          "(+ interpolation %.2f s)" % (t_host, t_host_all - t_host, t_dev, t_dev_all - t_dev),
          flush=True)
    assert np.array_equal(rho_h, rho_d)
    for a, b in zip(fh, fd):
        assert np.array_equal(a.getOctants(), b.getOctants())
        assert np.array_equal(a.getMeshConn(), b.getMeshConn())
        for x, y in zip(a.getDepNodeConn(), b.getDepNodeConn()):
            assert np.array_equal(x, y)
        assert np.array_equal(a.getPoints(), b.getPoints())
    for a, b in zip(ph, pd):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_box_balance_timing():
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names=workloads.FACES)
    f.createTrees(7)
    g = f.duplicate()
    t = time.time()
    f.balance(1)
    th = time.time() - t
    ms = g.balance(1, device=0)
    print("box 256x128x128 balance(1): host %.0f ms, device %.1f ms" % (1e3 * th, ms), flush=True)
    assert np.array_equal(f.getOctants(), g.getOctants())


def test_unbalanced_forest_is_refused():
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names=FACES)
    f.createTrees(2)
    f.refine(np.ones(len(f.getOctants()), np.intc), 0, 4)  # not balanced yet
    with pytest.raises(RuntimeError, match="tile the blocks"):
        f.createNodes(device=0)
