"""Coarse spaces that are not nested (CPU): per-block tree depths, the uneven
cube batch and the geometric interpolation between forests whose blocks
differ (tmrx_forest_create_interpolation_geometric).

The geometric interpolation has no counterpart in the reference; it is
pinned by (a) equality with createInterpolation (src/TMROctForest.cpp:
6611-6793, itself bit-exact against the oracle) wherever the pair is a
coarsen() pair, hanging nodes included, and (b) on pairs that are not
nested, by what trilinear interpolation fixes: rows sum to one, and every
global trilinear field (1, x, y, z, xy, yz, xz, xyz) is reproduced exactly,
because on box elements such a field lies in the coarse space."""
import numpy as np
import pytest
import scipy.sparse as sp

from tmr_amd import TMR, workloads


def _P(fine, coarse):
    rp, cl, w = fine.createGeometricInterpolation(coarse)
    return sp.csr_matrix((w, cl, rp), shape=(len(rp) - 1, int(cl.max()) + 1 if len(cl) else 0)), rp, cl, w


def _tri(X, c):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z
            + c[7] * x * y * z)


def test_split_and_levels():
    assert workloads.odd_split(48) == (3, 4)
    assert workloads.odd_split(13) == (13, 0)
    assert workloads.odd_split(8) == (1, 3)
    assert workloads.level_edges(13, 4) == [13, 7, 4, 2]
    assert workloads.level_edges(48, 4) == [48, 24, 12, 6]
    e = workloads.c5b_edges(96, 0)
    assert e.min() >= 8 and e.max() <= 48
    # the draw covers odd edges and edges with no power-of-two factor above 1
    assert any(int(n) % 2 for n in e) and len(set(int(n) for n in e)) > 20


def test_create_trees_per_block():
    conn, X = TMR.box_connectivity(3, 1, 2)
    f = TMR.OctForest()
    f.setConnectivity(conn)
    f.setNodePoints(X)
    lev = np.array([0, 2, 1, 3, 0, 1], dtype=np.intc)
    f.createTrees(block_levels=lev)
    o = f.getOctants()
    assert len(o) == int(np.sum(8 ** lev.astype(np.int64)))
    for b in range(6):
        ob = o[o["block"] == b]
        assert np.all(ob["level"] == lev[b])
        h = 1 << (30 - lev[b])
        cells = set(zip(ob["x"] // h, ob["y"] // h, ob["z"] // h))
        assert len(cells) == 8 ** lev[b]
    # every block at one depth: the same octants as createTrees(depth)
    g = TMR.OctForest()
    g.setConnectivity(conn)
    g.setNodePoints(X)
    g.createTrees(2)
    f.createTrees(block_levels=np.full(6, 2, dtype=np.intc))
    assert np.array_equal(f.getOctants(), g.getOctants())
    with pytest.raises(Exception):
        f.createTrees(block_levels=np.zeros(5, dtype=np.intc))


@pytest.mark.parametrize("which", ["uniform", "adaptive"])
def test_geometric_equals_create_interpolation_on_nested_pairs(which):
    if which == "uniform":
        forests, _ = workloads.uneven_cube_batch([8, 16, 12, 24], nlevels=3)
    else:
        forests, _, _ = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=3)
    for l in range(len(forests) - 1):
        a = forests[l].createInterpolation(forests[l + 1])
        b = forests[l].createGeometricInterpolation(forests[l + 1])
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        err = float(np.abs(a[2] - b[2]).max())
        print("%s level %d: %d rows, max weight difference %.1e" % (which, l, len(a[0]) - 1, err))
        assert err < 1e-13


def test_geometric_interpolation_not_nested():
    edges = [13, 9, 11, 24, 17]
    forests, rho = workloads.uneven_cube_batch(edges, nlevels=4)
    rng = np.random.default_rng(5)
    for l in range(3):
        f, c = forests[l], forests[l + 1]
        P, rp, cl, w = _P(f, c)
        Xf, Xc = f.getPoints(), c.getPoints()
        assert P.shape[0] == len(Xf) and int(cl.max()) < len(Xc)
        assert np.diff(rp).max() <= 8 and np.all(w > 0)
        assert np.abs(np.add.reduceat(w, rp[:-1]) - 1.0).max() < 1e-14
        for _ in range(3):
            coef = rng.standard_normal(8)
            err = np.abs(P @ _tri(Xc, coef) - _tri(Xf, coef)).max()
            assert err < 1e-12 * (1 + np.abs(_tri(Xf, coef)).max())
        # a field that is not trilinear is not reproduced (the check has teeth)
        assert np.abs(P @ Xc[:, 0] ** 2 - Xf[:, 0] ** 2).max() > 1e-6
    # cube k's nodes only take coarse nodes of cube k (cubes are apart)
    P, rp, cl, w = _P(forests[0], forests[1])
    xf = forests[0].getPoints()[:, 0]
    xc = forests[1].getPoints()[:, 0]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    assert np.abs(xf[rows] - xc[cl]).max() < 2 * 0.125 * 2


def test_uneven_batch_densities_and_names():
    edges = [9, 12]
    forests, rho = workloads.uneven_cube_batch(edges, nlevels=2, seed=4)
    assert len(rho) == 9 ** 3 + 12 ** 3
    # cube 0's densities are the seeded grid of that cube, element by element
    grid = workloads.smoothed_density((9, 9, 9), 4, passes=2)
    assert np.allclose(np.sort(rho[:729]), np.sort(grid.ravel()))
    for f in forests:
        assert len(f.getOctsWithName("fixed")) > 0 and len(f.getOctsWithName("load")) > 0
