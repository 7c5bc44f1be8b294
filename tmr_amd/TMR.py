"""Python mirror of the reference's TMR module for the hex8 solve path.

Class and method names, argument order and return types follow the Cython
binding tmr/TMR.pyx (OctForest :3393-3790, BoundaryConditions, createMg
:4317-4368, StiffnessProperties :5041-5184, TopoProblem :5371-6000).  MPI
communicators are accepted and ignored: one process owns the whole problem.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

UNIFORM_POINTS = 0
GAUSS_LOBATTO_POINTS = 1
BERNSTEIN_POINTS = 2
MAX_LEVEL = 30


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# TMROctant's layout (src/TMROctant.h:48-53)
OCT_DTYPE = np.dtype([("block", np.int32), ("x", np.int32), ("y", np.int32), ("z", np.int32),
                      ("tag", np.int32), ("level", np.int16), ("info", np.int16)])


def _octants_to_array(ptr, n):
    if n == 0:
        return np.zeros(0, dtype=OCT_DTYPE)
    buf = (C.c_char * (24 * n)).from_address(C.addressof(ptr.contents))
    return np.frombuffer(bytes(buf), dtype=OCT_DTYPE).copy()


class OctForest:
    """TMR.OctForest (tmr/TMR.pyx:3393-3790) on the native forest."""

    def __init__(self, comm=None, order=2, interp=GAUSS_LOBATTO_POINTS, _handle=None):
        self.comm = comm
        if _handle is not None:
            self.ptr = _handle
        else:
            h = C.c_void_p()
            check(lib().tmrx_forest_create(order, interp, C.byref(h)), "OctForest")
            self.ptr = h
        self._order = order
        self._interp = interp
        self._keep = []
        self._names = []  # (kind, index, name) as set, for save()

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().tmrx_forest_free(self.ptr)
                self.ptr = None
        except Exception:
            pass

    # -- setup --------------------------------------------------------------
    def getMeshOrder(self):
        return self._order

    def getInterpType(self):
        return self._interp

    def setMeshOrder(self, order, interp=GAUSS_LOBATTO_POINTS):
        """TMROctForest::setMeshOrder: frees the nodes; only order 2 is built."""
        check(lib().tmrx_forest_set_mesh_order(self.ptr, int(order), int(interp)), "setMeshOrder")
        self._order, self._interp = max(int(order), 2), interp

    def setConnectivity(self, conn):
        conn = np.ascontiguousarray(conn, dtype=np.intc)
        if conn.ndim != 2 or conn.shape[1] != 8:
            raise ValueError("Expected array of shape (n, 8)")
        num_nodes = int(conn.max()) + 1
        check(lib().tmrx_forest_set_connectivity(self.ptr, num_nodes, _iptr(conn), conn.shape[0]),
              "setConnectivity")

    def setNodePoints(self, X):
        """Block vertex coordinates (num_nodes x 3); stands in for setTopology."""
        X = np.ascontiguousarray(X, dtype=np.double)
        check(lib().tmrx_forest_set_node_points(self.ptr, _dptr(X)), "setNodePoints")
        self._Xv = X.copy()

    def getElementExtents(self):
        """Edge lengths (nelem x 3) of every element of a forest of box blocks."""
        octs = self.getOctants()
        bc = self.getConnectivity()["block_conn"]
        X = self._Xv
        ext = X[bc[:, 7]] - X[bc[:, 0]]
        return ext[octs["block"]] * np.ldexp(1.0, -octs["level"].astype(np.int64))[:, None]

    def setName(self, kind, index, name):
        kinds = {"vertex": 0, "edge": 1, "face": 2, "volume": 3, "block": 3}
        k = kinds[kind] if isinstance(kind, str) else int(kind)
        check(lib().tmrx_forest_set_name(self.ptr, k, int(index), name.encode()), "setName")
        self._names.append((k, int(index), str(name)))

    def getConnectivity(self):
        nb, nf, ne, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        bc, bfc, bec, bfi = (C.POINTER(C.c_int)() for _ in range(4))
        check(lib().tmrx_forest_get_connectivity(self.ptr, C.byref(nb), C.byref(nf), C.byref(ne),
                                                 C.byref(nn), C.byref(bc), C.byref(bfc),
                                                 C.byref(bec), C.byref(bfi)), "getConnectivity")
        b = nb.value
        return dict(nblocks=b, nfaces=nf.value, nedges=ne.value, nnodes=nn.value,
                    block_conn=np.ctypeslib.as_array(bc, (8 * b,)).reshape(b, 8).copy(),
                    block_face_conn=np.ctypeslib.as_array(bfc, (6 * b,)).reshape(b, 6).copy(),
                    block_edge_conn=np.ctypeslib.as_array(bec, (12 * b,)).reshape(b, 12).copy(),
                    block_face_ids=np.ctypeslib.as_array(bfi, (6 * b,)).reshape(b, 6).copy())

    def repartition(self, max_rank=-1):
        check(lib().tmrx_forest_repartition(self.ptr), "repartition")

    def setOctants(self, octs):
        """The octants of a saved forest (getOctants' structured array):
        validated (inside their blocks, aligned, tiling every block) and
        sorted into Morton order; frees the nodes."""
        a = np.ascontiguousarray(octs, dtype=OCT_DTYPE)
        check(lib().tmrx_forest_set_octants(self.ptr, C.c_void_p(a.ctypes.data), len(a)),
              "setOctants")

    def save(self, fname):
        """Checkpoint of the forest (block connectivity and vertices, names,
        mesh order, octants) as an .npz file; loadForest rebuilds it."""
        names = np.array([n for _, _, n in self._names], dtype=str)
        kinds = np.array([(k, i) for k, i, _ in self._names], dtype=np.int64).reshape(-1, 2)
        np.savez(fname, block_conn=self.getConnectivity()["block_conn"],
                 points=getattr(self, "_Xv", np.zeros((0, 3))), name_kinds=kinds,
                 names=names, octants=self.getOctants(),
                 order=np.array([self._order, self._interp]))

    def writeForestToVTK(self, fname):
        """TMROctForest::writeForestToVTK (src/TMROctForest.cpp:1317): the
        octants as hexahedra (legacy VTK), no fields."""
        writeVTK(fname, self)

    def createTrees(self, depth=0, block_levels=None):
        """Uniform trees of `depth`; with `block_levels` (one per block) each
        block gets its own depth (blocks of different sizes, elements of one
        size)."""
        if block_levels is None:
            check(lib().tmrx_forest_create_trees(self.ptr, int(depth)), "createTrees")
            return
        lv = np.ascontiguousarray(block_levels, dtype=np.intc)
        check(lib().tmrx_forest_create_trees_per_block(self.ptr, _iptr(lv), len(lv)), "createTrees")

    def refine(self, refine=None, min_lev=0, max_lev=MAX_LEVEL):
        if refine is not None:
            refine = np.ascontiguousarray(refine, dtype=np.intc)
            n = len(self.getOctants())
            if refine.shape != (n,):
                raise ValueError("refine must have one entry per octant")
            check(lib().tmrx_forest_refine(self.ptr, _iptr(refine), int(min_lev), int(max_lev)), "refine")
        else:
            check(lib().tmrx_forest_refine(self.ptr, None, int(min_lev), int(max_lev)), "refine")

    def duplicate(self):
        h = C.c_void_p()
        check(lib().tmrx_forest_duplicate(self.ptr, C.byref(h)), "duplicate")
        f = OctForest(self.comm, self._order, self._interp, _handle=h)
        f._Xv = getattr(self, "_Xv", None)
        f._names = list(self._names)
        return f

    def coarsen(self):
        h = C.c_void_p()
        check(lib().tmrx_forest_coarsen(self.ptr, C.byref(h)), "coarsen")
        f = OctForest(self.comm, self._order, self._interp, _handle=h)
        f._Xv = getattr(self, "_Xv", None)
        f._names = list(self._names)
        return f

    def balance(self, btype, device=None):
        """TMROctForest::balance; with `device` (a GPU index) on that GPU.
        Returns the device pass's wall time in ms, or None on the host."""
        if device is None:
            check(lib().tmrx_forest_balance(self.ptr, int(btype)), "balance")
            return None
        ms = C.c_double()
        check(lib().tmrx_forest_balance_device(self.ptr, int(btype), int(device), C.byref(ms)),
              "balance")
        return ms.value

    def createNodes(self, device=None):
        """TMROctForest::createNodes; with `device` (a GPU index) the same
        nodes are built on that GPU.  Returns the device pass's wall time in
        ms, or None on the host."""
        if device is None:
            check(lib().tmrx_forest_create_nodes(self.ptr), "createNodes")
            return None
        ms = C.c_double()
        check(lib().tmrx_forest_create_nodes_device(self.ptr, int(device), C.byref(ms)),
              "createNodes")
        return ms.value

    # -- queries --------------------------------------------------------------
    def getOctants(self):
        p = C.POINTER(_lib.Octant)()
        n = C.c_int()
        check(lib().tmrx_forest_get_octants(self.ptr, C.byref(p), C.byref(n)), "getOctants")
        return _octants_to_array(p, n.value)

    def getOctsWithName(self, aname):
        name = aname.encode() if aname is not None else None
        n = check(lib().tmrx_forest_get_octs_with_name(self.ptr, name, None, 0), "getOctsWithName")
        buf = (_lib.Octant * max(n, 1))()
        check(lib().tmrx_forest_get_octs_with_name(self.ptr, name, buf, n), "getOctsWithName")
        return _octants_to_array(C.cast(buf, C.POINTER(_lib.Octant)), n)

    def getNodesWithName(self, aname):
        name = aname.encode() if aname is not None else None
        n = check(lib().tmrx_forest_get_nodes_with_name(self.ptr, name, None, 0), "getNodesWithName")
        out = np.zeros(max(n, 1), dtype=np.intc)
        check(lib().tmrx_forest_get_nodes_with_name(self.ptr, name, _iptr(out), n), "getNodesWithName")
        return out[:n].copy()

    def getNodeRange(self):
        r = np.zeros(2, dtype=np.intc)
        check(lib().tmrx_forest_get_owned_node_range(self.ptr, _iptr(r)), "getNodeRange")
        return r

    def getMeshConn(self):
        p = C.POINTER(C.c_int)()
        ne, no = C.c_int(), C.c_int()
        check(lib().tmrx_forest_get_node_conn(self.ptr, C.byref(p), C.byref(ne), C.byref(no),
                                              None), "getMeshConn")
        if ne.value == 0:
            return np.zeros((0, 8), dtype=np.intc)
        return np.ctypeslib.as_array(p, (8 * ne.value,)).reshape(ne.value, 8).copy()

    def getDepNodeConn(self):
        p, c = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        w = C.POINTER(C.c_double)()
        nd = check(lib().tmrx_forest_get_dep_node_conn(self.ptr, C.byref(p), C.byref(c), C.byref(w)),
                   "getDepNodeConn")
        ptr = np.ctypeslib.as_array(p, (nd + 1,)).copy()
        nnz = int(ptr[-1])
        if nnz == 0:
            return ptr, np.zeros(0, np.intc), np.zeros(0)
        return (ptr, np.ctypeslib.as_array(c, (nnz,)).copy(),
                np.ctypeslib.as_array(w, (nnz,)).copy())

    def getNodeNumbers(self):
        p = C.POINTER(C.c_int)()
        n = C.c_int()
        check(lib().tmrx_forest_get_node_numbers(self.ptr, C.byref(p), C.byref(n)), "getNodeNumbers")
        return np.ctypeslib.as_array(p, (n.value,)).copy()

    def getPoints(self):
        p = C.POINTER(C.c_double)()
        n = C.c_int()
        check(lib().tmrx_forest_get_points(self.ptr, C.byref(p), C.byref(n)), "getPoints")
        return np.ctypeslib.as_array(p, (3 * n.value,)).reshape(n.value, 3).copy()

    def createInterpolation(self, forest, vec=None, device=None):
        """CSR of the prolongation from `forest` (coarse) to this forest.

        Returns (rowp, cols, weights); if `vec` is given (an object with a
        `setCSR(rowp, cols, weights, ncols)` method) it is filled instead.
        With `device` (a GPU index) the same CSR is built on that GPU.
        """
        if forest.ptr.value == self.ptr.value:
            raise ValueError("Cannot interpolate between the same object")
        nr, nc = C.c_int(), C.c_int()
        rp, cl = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        w = C.POINTER(C.c_double)()
        if device is None:
            st = lib().tmrx_forest_create_interpolation(self.ptr, forest.ptr, C.byref(nr),
                                                         C.byref(nc), C.byref(rp), C.byref(cl),
                                                         C.byref(w))
        else:
            self.last_device_ms = C.c_double()
            st = lib().tmrx_forest_create_interpolation_device(
                self.ptr, forest.ptr, int(device), C.byref(nr), C.byref(nc), C.byref(rp),
                C.byref(cl), C.byref(w), C.byref(self.last_device_ms))
            self.last_device_ms = self.last_device_ms.value
        check(st, "createInterpolation")
        return self._csr_out(nr, nc, rp, cl, w, vec)

    def createGeometricInterpolation(self, forest, vec=None):
        """Interpolation from `forest` (coarse) to this forest by physical
        position, for forests whose blocks differ (a coarse space that is not
        a coarsen() of this forest).  Same return convention as
        createInterpolation; on a coarsen() pair the two agree to 1e-13."""
        if forest.ptr.value == self.ptr.value:
            raise ValueError("Cannot interpolate between the same object")
        nr, nc = C.c_int(), C.c_int()
        rp, cl = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        w = C.POINTER(C.c_double)()
        check(lib().tmrx_forest_create_interpolation_geometric(
            self.ptr, forest.ptr, C.byref(nr), C.byref(nc), C.byref(rp), C.byref(cl),
            C.byref(w)), "createGeometricInterpolation")
        return self._csr_out(nr, nc, rp, cl, w, vec)

    @staticmethod
    def _csr_out(nr, nc, rp, cl, w, vec):
        rowp = np.ctypeslib.as_array(rp, (nr.value + 1,)).copy()
        nnz = int(rowp[-1])
        cols = np.ctypeslib.as_array(cl, (nnz,)).copy() if nnz else np.zeros(0, np.intc)
        wts = np.ctypeslib.as_array(w, (nnz,)).copy() if nnz else np.zeros(0)
        if vec is not None:
            vec.setCSR(rowp, cols, wts, nc.value)
            return vec
        return rowp, cols, wts

    def createElementCenterWeights(self, filt):
        """Element-centre density weights from the filter forest `filt`
        (TMROctTACSTopoCreator::computeWeights, src/topology/
        TMR_TACSTopoCreator.cpp:55-186): returns CSR (rowp, cols, weights)
        with one row per element of this forest and columns = independent
        filter nodes, so that rho_e = sum_j w_ej x_j."""
        nr, nc = C.c_int(), C.c_int()
        rp, cl = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        w = C.POINTER(C.c_double)()
        check(lib().tmrx_forest_element_center_weights(self.ptr, filt.ptr, C.byref(nr), C.byref(nc),
                                                       C.byref(rp), C.byref(cl), C.byref(w)),
              "createElementCenterWeights")
        rowp = np.ctypeslib.as_array(rp, (nr.value + 1,)).copy()
        nnz = int(rowp[-1])
        cols = np.ctypeslib.as_array(cl, (nnz,)).copy() if nnz else np.zeros(0, np.intc)
        wts = np.ctypeslib.as_array(w, (nnz,)).copy() if nnz else np.zeros(0)
        return rowp, cols, wts, nc.value


def loadForest(fname, comm=None):
    """The forest saved by OctForest.save: the same blocks, names, mesh order
    and octants (nodes are created on demand, as for any forest)."""
    d = np.load(fname, allow_pickle=False)
    order, interp = (int(v) for v in d["order"])
    f = OctForest(comm, order=order, interp=interp)
    f.setConnectivity(d["block_conn"])
    if len(d["points"]):
        f.setNodePoints(d["points"])
    for (k, i), n in zip(d["name_kinds"], d["names"]):
        f.setName(int(k), int(i), str(n))
    f.setOctants(d["octants"])
    return f


def box_connectivity(nx, ny, nz, L=(1.0, 1.0, 1.0)):
    """Block connectivity and vertex points of an nx x ny x nz lattice of box
    blocks covering [0,Lx]x[0,Ly]x[0,Lz]; corner ordering i + 2j + 4k."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()  # block i + nx (j + ny k)
    vid = lambda a, b, c: (i + a) + (nx + 1) * ((j + b) + (ny + 1) * (k + c))
    conn = np.stack([vid(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)], axis=1)
    vk, vj, vi = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    X = np.stack([L[0] * vi.ravel() / nx, L[1] * vj.ravel() / ny, L[2] * vk.ravel() / nz], axis=1)
    return np.ascontiguousarray(conn, dtype=np.intc), np.ascontiguousarray(X, dtype=float)


def createBoxForest(nx, ny, nz, L=(1.0, 1.0, 1.0), face_names=None, comm=None):
    """Forest on an nx x ny x nz lattice of box blocks with named outer faces.

    face_names maps 'x-', 'x+', 'y-', 'y+', 'z-', 'z+' to names; every block
    face on that side of the box receives it.  This replaces the CAD topology
    (TMRTopology) the reference uses to attach names.
    """
    conn, X = box_connectivity(nx, ny, nz, L)
    f = OctForest(comm)
    f.setConnectivity(conn)
    f.setNodePoints(X)
    if face_names:
        info = f.getConnectivity()
        bfc = info["block_face_conn"]
        sides = {"x-": 0, "x+": 1, "y-": 2, "y+": 3, "z-": 4, "z+": 5}
        dims = (nx, ny, nz)
        for side, name in face_names.items():
            fi = sides[side]
            a, hi = fi // 2, fi % 2
            for b in range(len(conn)):
                ijk = (b % nx, (b // nx) % ny, b // (nx * ny))
                if ijk[a] == (dims[a] - 1 if hi else 0):
                    f.setName("face", int(bfc[b, fi]), name)
    return f


def writeSTLToBin(fname, forest, x, index=0, cutoff=0.5, vars_per_node=None):
    """TMR.writeSTLToBin (tmr/TMR.pyx:4610-4631, TMR_GenerateBinFile): the
    surface of {x >= cutoff} of a nodal design field as a binary triangle
    file; returns the triangle count.  x holds vars_per_node values per
    independent node (inferred from its length), component `index`."""
    x = np.ascontiguousarray(x, dtype=np.double)
    n = int(forest.getNodeRange()[1])
    if vars_per_node is None:
        if n == 0 or x.size % n:
            raise ValueError("x must hold a whole number of values per node")
        vars_per_node = x.size // n
    nt = C.c_int()
    check(lib().tmrx_forest_generate_bin_file(forest.ptr, str(fname).encode(),
                                              x.ctypes.data_as(C.POINTER(C.c_double)),
                                              int(vars_per_node), int(index), float(cutoff),
                                              C.byref(nt)), "writeSTLToBin")
    return nt.value


def convertBinToSTL(binfile, stlfile):
    """TMR_ConvertBinToSTL: the binary triangle file as ASCII STL."""
    check(lib().tmrx_convert_bin_to_stl(str(binfile).encode(), str(stlfile).encode()),
          "convertBinToSTL")


def writeVTK(fname, forest, node_field=None, node_name="u", elem_field=None, elem_name="rho"):
    """Volume output of the forest's mesh (the role of TACS's ToFH5 .f5 file
    in the reference's examples): legacy-VTK hex8 grid with an optional
    nodal field (values per independent node, any number of components;
    hanging nodes interpolated) and an optional element field."""
    n, ne = int(forest.getNodeRange()[1]), len(forest.getOctants())
    nv = ev = 0
    nptr = eptr = None
    if node_field is not None:
        node_field = np.ascontiguousarray(node_field, dtype=np.double).reshape(-1)
        if n == 0 or node_field.size % n:
            raise ValueError("node_field must hold a whole number of values per node")
        nv, nptr = node_field.size // n, node_field.ctypes.data
    if elem_field is not None:
        elem_field = np.ascontiguousarray(elem_field, dtype=np.double).reshape(-1)
        if ne == 0 or elem_field.size % ne:
            raise ValueError("elem_field must hold a whole number of values per element")
        ev, eptr = elem_field.size // ne, elem_field.ctypes.data
    check(lib().tmrx_forest_write_vtk(forest.ptr, str(fname).encode(), nptr, nv,
                                      node_name.encode(), eptr, ev, elem_name.encode()),
          "writeVTK")


def readSTLBin(binfile):
    """The triangles of a binary triangle file, shape (ntris, 3, 3)."""
    with open(binfile, "rb") as fh:
        nt = int(np.frombuffer(fh.read(4), dtype=np.int32)[0])
        return np.frombuffer(fh.read(72 * nt), dtype=np.double).reshape(nt, 3, 3).copy()


# topology-optimisation classes (tmr/TMR.pyx:3877-6005) live in topo.py
from .topo import (BoundaryConditions, Compliance, ConformFilter, LagrangeFilter,  # noqa: E402,F401
                   HelmholtzFilter, HelmholtzPUFilter, MatrixFilter,
                   MaterialProperties, OctConformTopoCreator, OctConstitutive, OctTopoCreator,
                   PNormFailure, StiffnessProperties, StructuralMass, TopoProblem, createMg,
                   RAMP_PENALTY, SIMP_PENALTY)
