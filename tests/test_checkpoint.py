"""Forest checkpoints (CPU): OctForest.save / TMR.loadForest rebuild the same
forest -- blocks, names, octants -- so the node numbering, hanging-node
constraints, boundary nodes and interpolation of the reloaded forest equal
the original's bit for bit; tmrx_forest_set_octants refuses octant sets
that do not tile their blocks."""
import numpy as np
import pytest

from tmr_amd import TMR, workloads


def _maps(f):
    f.createNodes()
    ptr, conn, w = f.getDepNodeConn()
    return (np.asarray(f.getMeshConn()), np.asarray(ptr), np.asarray(conn), np.asarray(w),
            np.asarray(f.getNodesWithName("fixed")), np.asarray(f.getPoints()))


def test_save_load_adaptive_hierarchy(tmp_path):
    forests, rho, L = workloads.adaptive(seed=3, base_level=2, max_level=4, nlevels=3)
    loaded = []
    for l, f in enumerate(forests):
        p = str(tmp_path / ("level%d.npz" % l))
        f.save(p)
        g = TMR.loadForest(p)
        a, b = _maps(f), _maps(g)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.array_equal(f.getOctants(), g.getOctants())
        loaded.append(g)
    # the reloaded hierarchy interpolates exactly as the original one
    for l in range(len(forests) - 1):
        pa = forests[l].createInterpolation(forests[l + 1])
        pb = loaded[l].createInterpolation(loaded[l + 1])
        for x, y in zip(pa, pb):
            assert np.array_equal(x, y)


def test_save_load_per_block_depths(tmp_path):
    forests, rho = workloads.uneven_cube_batch([9, 12], nlevels=2)
    p = str(tmp_path / "cubes.npz")
    forests[0].save(p)
    g = TMR.loadForest(p)
    for x, y in zip(_maps(forests[0]), _maps(g)):
        assert np.array_equal(x, y)
    assert len(g.getOctsWithName("load")) == len(forests[0].getOctsWithName("load"))


def test_set_octants_rejects_bad_sets():
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names={"x-": "fixed"})
    f.createTrees(2)
    good = f.getOctants()
    g = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names={"x-": "fixed"})
    g.setOctants(good[::-1])  # any order: sorted on the way in
    assert np.array_equal(g.getOctants(), good)
    bad = [good[1:],                                   # a gap
           np.concatenate([good, good[:1]])]           # an overlap
    o = good.copy()
    o["x"][5] += 1                                     # not aligned to its level
    bad.append(o)
    o = good.copy()
    o["block"][0] = 7                                  # no such block
    bad.append(o)
    o = good.copy()
    o["level"][0] = 3                                  # a finer octant leaves a gap
    bad.append(o)
    for b in bad:
        with pytest.raises(Exception):
            g.setOctants(b)
    # a coarser octant covering its siblings' space tiles the block again
    h = TMR.createBoxForest(1, 1, 1, face_names={"x-": "fixed"})
    h.createTrees(1)
    o = h.getOctants()[:1].copy()
    o["level"][0] = 0
    h.setOctants(o)
    assert len(h.getOctants()) == 1


def test_write_forest_to_vtk(tmp_path):
    """writeForestToVTK (TMROctForest::writeForestToVTK): one hexahedron per
    octant over the forest's nodes, no fields."""
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names={"x-": "fixed"})
    f.createTrees(2)
    f.createNodes()
    p = str(tmp_path / "forest.vtk")
    f.writeForestToVTK(p)
    lines = open(p).read().splitlines()
    head = {l.split()[0]: l.split()[1:] for l in lines if l[:1].isupper() and " " in l}
    assert int(head["POINTS"][0]) == 9 * 5 * 5
    assert int(head["CELLS"][0]) == len(f.getOctants()) == 128
    assert "POINT_DATA" not in head and "CELL_DATA" not in head
