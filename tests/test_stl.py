"""Surface output (TMR_GenerateBinFile / TMR_ConvertBinToSTL) pinned by
closed forms: for a field linear in space the marching-tetrahedra level set
is exact, so the surface of {x >= cutoff} has the exact area and encloses the
exact volume (divergence theorem), is closed (every edge shared by two
triangles with opposite orientations) on a conforming mesh, and survives the
ASCII conversion."""
import numpy as np
import pytest

from tmr_amd import TMR, workloads


def _points(f):
    nums, X = f.getNodeNumbers(), f.getPoints()
    P = np.zeros((int(f.getNodeRange()[1]), 3))
    P[nums[nums >= 0]] = X[nums >= 0]
    return P


def _area_volume(T):
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    area = 0.5 * np.linalg.norm(n, axis=1).sum()
    vol = np.einsum("ij,ij->i", T[:, 0], n).sum() / 6.0
    return area, vol


def test_oblique_plane_in_unit_cube(tmp_path):
    """{(x + y + z) / 3 >= 0.5} in the unit cube: volume 1/2 by symmetry;
    area = the hexagonal section 3 sqrt(3) / 4 plus 7/8 of each face at 1
    and 1/8 of each face at 0 (3 in all)."""
    f = TMR.createBoxForest(1, 1, 1)
    f.createTrees(3)
    f.createNodes()
    P = _points(f)
    x = P.sum(1) / 3.0
    nt = TMR.writeSTLToBin(tmp_path / "s.bin", f, x, cutoff=0.5)
    T = TMR.readSTLBin(tmp_path / "s.bin")
    assert nt == len(T) > 0
    area, vol = _area_volume(T)
    assert abs(vol - 0.5) < 1e-12
    assert abs(area - (3.0 * np.sqrt(3.0) / 4.0 + 3.0)) < 1e-12
    # closed: every directed edge has its reverse
    edges = {}
    for t in np.round(T, 12):
        for k in range(3):
            a, b = tuple(t[k]), tuple(t[(k + 1) % 3])
            edges[(a, b)] = edges.get((a, b), 0) + 1
    assert all(edges.get((b, a), 0) == c for (a, b), c in edges.items())
    TMR.convertBinToSTL(tmp_path / "s.bin", tmp_path / "s.stl")
    txt = (tmp_path / "s.stl").read_text().split("\n")
    assert txt[0] == "solid topology" and txt[-2] == "endsolid topology"
    assert sum(1 for line in txt if line.startswith("facet normal")) == nt
    v = np.array([[float(w) for w in line.split()[1:]] for line in txt if line.startswith("vertex")])
    assert np.allclose(v.reshape(-1, 3, 3), T, rtol=1e-6, atol=1e-7)


def test_axis_plane_with_hanging_nodes(tmp_path):
    """{x >= 1} of x / 2 on the adaptive 2 x 1 x 1 forest (hanging nodes take
    the exact linear value): the unit cube [1, 2] x [0, 1] x [0, 1], area 6,
    volume 1; a second value per node selects the other half."""
    forests, rho, L = workloads.adaptive(seed=2, base_level=2, max_level=4, nlevels=1)
    f = forests[0]
    assert len(f.getDepNodeConn()[0]) > 1
    P = _points(f)
    x2 = np.stack([P[:, 0] / 2.0, 1.0 - P[:, 0] / 2.0], 1).reshape(-1)
    for idx in (0, 1):
        TMR.writeSTLToBin(tmp_path / "a.bin", f, x2, index=idx, cutoff=0.5)
        area, vol = _area_volume(TMR.readSTLBin(tmp_path / "a.bin"))
        assert abs(vol - 1.0) < 1e-12 and abs(area - 6.0) < 1e-12
    assert TMR.writeSTLToBin(tmp_path / "e.bin", f, np.zeros(len(P)), cutoff=0.5) == 0


def _read_vtk(path):
    """Minimal reader of the legacy-VTK files written by writeVTK."""
    tok = open(path).read().split()
    i = tok.index("POINTS")
    n = int(tok[i + 1])
    P = np.array(tok[i + 3:i + 3 + 3 * n], float).reshape(n, 3)
    i = tok.index("CELLS")
    ne = int(tok[i + 1])
    cells = np.array(tok[i + 3:i + 3 + 9 * ne], int).reshape(ne, 9)
    assert np.all(cells[:, 0] == 8)
    i = tok.index("CELL_TYPES")
    assert all(t == "12" for t in tok[i + 2:i + 2 + ne])
    out = {"P": P, "cells": cells[:, 1:]}
    if "VECTORS" in tok:
        i = tok.index("VECTORS")
        out["u"] = np.array(tok[i + 3:i + 3 + 3 * n], float).reshape(n, 3)
    if "CELL_DATA" in tok:
        i = tok.index("CELL_DATA")
        out["rho"] = np.array(tok[i + 8:i + 8 + ne], float)
    return out


def _hex_volume(V):
    """Volume of a hexahedron in VTK corner order (five tetrahedra)."""
    tets = [(0, 1, 3, 4), (1, 2, 3, 6), (1, 4, 5, 6), (3, 4, 6, 7), (1, 3, 4, 6)]
    vol = 0.0
    for a, b, c, d in tets:
        vol += np.dot(V[b] - V[a], np.cross(V[c] - V[a], V[d] - V[a])) / 6.0
    return vol


def test_vtk_volume_output(tmp_path):
    """writeVTK: every hexahedron in VTK order has the element's volume
    (positive orientation), hanging points carry the interpolated nodal
    field (exact for a linear field), element data round trip."""
    f = TMR.createBoxForest(2, 1, 1, (2.0, 1.0, 1.0), face_names={"x-": "fixed"})
    f.createTrees(1)
    rng = np.random.default_rng(3)
    f.refine(rng.integers(0, 2, len(f.getOctants())).astype(np.intc), 0, 3)
    f.balance(1)
    f.createNodes()
    nums, X = f.getNodeNumbers(), f.getPoints()
    n = int(f.getNodeRange()[1])
    Pind = np.zeros((n, 3))
    Pind[nums[nums >= 0]] = X[nums >= 0]
    A = np.array([[0.5, -1.0, 2.0], [1.5, 0.25, -0.5], [0.0, 1.0, 1.0]])
    u = Pind @ A.T + np.array([0.1, 0.2, 0.3])
    rho = rng.random(len(f.getOctants()))
    TMR.writeVTK(tmp_path / "m.vtk", f, node_field=u, elem_field=rho)
    d = _read_vtk(tmp_path / "m.vtk")
    assert len(d["P"]) == len(nums) and np.allclose(d["P"], X, rtol=0, atol=1e-15)
    ext = f.getElementExtents()
    for e, c in enumerate(d["cells"]):
        assert abs(_hex_volume(d["P"][c]) - np.prod(ext[e])) < 1e-12
    assert np.allclose(d["u"], d["P"] @ A.T + np.array([0.1, 0.2, 0.3]), atol=1e-13)
    assert np.array_equal(d["rho"], rho)
    assert (nums < 0).any()
