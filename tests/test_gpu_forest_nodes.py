"""createNodes on the GPU (tmrx_forest_create_nodes_device) against the host
pass (tmrx_forest_create_nodes, itself bit-exact with the geometric oracle
in test_forest.py): octant info bits, element connectivity, hanging-node
constraints, node numbers and points equal bit for bit on uniform, adaptive
(face- and corner-balanced), eight-block and rotated-block forests, and on
the 1.2M-hex adaptive and 4.2M-hex box forests, where the two passes are
timed side by side."""
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


def test_unbalanced_forest_is_refused():
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names=FACES)
    f.createTrees(2)
    f.refine(np.ones(len(f.getOctants()), np.intc), 0, 4)  # not balanced yet
    with pytest.raises(RuntimeError, match="tile the blocks"):
        f.createNodes(device=0)
