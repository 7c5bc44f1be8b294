"""TMR's TopOptUtils helpers (tmr/TopOptUtils.py:14-307, 397-457) over the
native problem: the level-hierarchy builder, the load vectors, the design
transfer between adapted meshes and the density-based refinement rule.

Loads are host-side setup: consistent nodal forces of a constant traction on
the faces of axis-aligned box elements (exact for bilinear faces: a quarter
of ``t * area`` at each face corner), hanging corners distributed to their
parents with the forest's dependent weights, clamped rows zeroed as TACS's
``assembleRes``/``applyBCs`` leaves them.
"""
import numpy as np

from . import TMR, topo


def createTopoProblem(forest, callback, filter_type, nlevels=2, repartition=True,
                      design_vars_per_node=1, r0=0.05, N=10, lowest_order=2,
                      ordering=topo.MULTICOLOR_ORDER, use_galerkin=False,
                      scale_coordinate_factor=1.0):
    """Hierarchy of forests (balance(1), coarsen per level), one creator and
    filter per level from ``creator, filter = callback(forest)``, the
    multigrid object and the TopoProblem (TopOptUtils.py:14-128).

    Only order-2 meshes.  Filters: "lagrange" (element densities
    interpolated from the filter forest, OctTopoCreator), "conform" (nodal
    design variables on the analysis forests, OctConformTopoCreator), and
    "matrix" / "helmholtz" (conform levels whose analysis design is the
    matrix- or Helmholtz-filtered design, radius r0, N terms), or a callable
    filter_type(assemblers, filters) returning the filter (e.g. an
    initialised Mfilter, as the reference's examples pass it)."""
    if design_vars_per_node != 1:
        raise NotImplementedError("one design variable per node only")
    if scale_coordinate_factor != 1.0:
        raise NotImplementedError("scale_coordinate_factor: set the block points instead")
    forests, filters, assemblers = [], [], []
    forest.balance(1)
    if repartition:
        forest.repartition()
    creator, filtr = callback(forest)
    forests.append(forest)
    filters.append(filtr)
    assemblers.append(creator.createTACS(forest, ordering))
    for _ in range(nlevels - 1):
        order = forests[-1].getMeshOrder()
        if order > lowest_order:
            raise NotImplementedError("mesh order %d: only order-2 meshes are supported" % order)
        forest = forests[-1].coarsen()
        forest.balance(1)
        if repartition:
            forest.repartition()
        creator, filtr = callback(forest)
        forests.append(forest)
        filters.append(filtr)
        assemblers.append(creator.createTACS(forest, ordering))
    mg = topo.createMg(assemblers, forests, use_galerkin=use_galerkin)
    if callable(filter_type):
        filter_obj = filter_type(assemblers, filters)
    elif filter_type == "lagrange":
        filter_obj = topo.LagrangeFilter(assemblers, filters)
    elif filter_type == "conform":
        filter_obj = topo.ConformFilter(assemblers, filters)
    elif filter_type == "matrix":
        filter_obj = topo.MatrixFilter(r0, N, assemblers, filters)
    elif filter_type == "helmholtz":
        filter_obj = topo.HelmholtzFilter(r0, assemblers, filters)
    else:
        raise NotImplementedError("filter type %r is not supported" % (filter_type,))
    return topo.TopoProblem(filter_obj, mg)


def _clamp(assembler, forest, f):
    """Zero the clamped rows (TACSAssembler::applyBCs)."""
    for name, mask in assembler.bcs.as_list():
        nodes = forest.getNodesWithName(name)
        nodes = nodes[nodes >= 0]
        for d in range(3):
            if mask & (1 << d):
                f[3 * nodes + d] = 0.0
    return f


def _add_node_force(f, node, F, dep):
    """Add the 3-vector F at a node; a hanging node passes it to its
    parents with its dependent weights (the transpose of the constraint)."""
    ptr, dconn, dw = dep
    if node >= 0:
        f[3 * node:3 * node + 3] += F
    else:
        d = -node - 1
        for j in range(ptr[d], ptr[d + 1]):
            f[3 * dconn[j]:3 * dconn[j] + 3] += dw[j] * F


def computeTractionLoad(names, forest, assembler, trac):
    """Load of a traction on every octant face carrying one of ``names``;
    ``trac[face]`` is the traction (3 components) for local face index
    ``face`` (0..5: x-, x+, y-, y+, z-, z+), as in TopOptUtils.py:176-238."""
    names = [names] if isinstance(names, str) else list(names)
    trac = [np.asarray(t, dtype=np.double) for t in trac]
    if len(trac) != 6 or any(t.shape != (3,) for t in trac):
        raise ValueError("expected one 3-component traction per octant face (6)")
    conn = forest.getMeshConn()
    ext = forest.getElementExtents()
    dep = forest.getDepNodeConn()
    f = assembler.createVec()
    for name in names:
        for o in forest.getOctsWithName(name):
            e, face = int(o["tag"]), int(o["info"])
            ax, side = face // 2, face % 2
            area = ext[e, (ax + 1) % 3] * ext[e, (ax + 2) % 3]
            F = 0.25 * area * trac[face]
            for k in range(8):
                if ((k >> ax) & 1) == side:
                    _add_node_force(f, int(conn[e, k]), F, dep)
    return _clamp(assembler, forest, f)


def compute3DTractionLoad(name, forest, assembler, tr):
    """Constant traction ``tr`` on the faces named ``name``
    (TopOptUtils.py:241-267)."""
    tr = np.asarray(tr, dtype=np.double)
    return computeTractionLoad(name, forest, assembler, [tr] * 6)


def computeVertexLoad(name, forest, assembler, point_force):
    """Point force at each independent node carrying ``name``
    (TopOptUtils.py:131-173: nodes outside the owned range -- here the
    hanging ones -- receive nothing)."""
    point_force = np.asarray(point_force, dtype=np.double)
    if len(point_force) != assembler.getVarsPerNode():
        raise ValueError("Point force length must be equal to vars_per_node")
    f = assembler.createVec()
    nodes = forest.getNodesWithName(name)
    lo, hi = forest.getNodeRange()
    for n in nodes:
        if lo <= n < hi:
            f[3 * (n - lo):3 * (n - lo) + 3] += point_force
    return f


def interpolateDesignVec(orig_filter, orig_vec, new_filter, new_vec=None):
    """Design variables of an adapted mesh (TopOptUtils.interpolateDesignVec,
    tmr/TopOptUtils.py:270-307): new_vec = P orig_vec with
    P = new_filter.createInterpolation(orig_filter), nodal values
    interpolated at the new forest's independent nodes; the product runs on
    the GPU (tmrx_forest_interpolate_vec).  Vectors are numpy arrays (TMR
    passes ParOpt vectors); several values per node are interleaved.
    Returns new_vec."""
    import ctypes as C
    from ._lib import check, lib
    orig_vec = np.ascontiguousarray(orig_vec, dtype=np.double)
    n_orig = int(orig_filter.getNodeRange()[1])
    n_new = int(new_filter.getNodeRange()[1])
    if n_orig == 0 or orig_vec.size % n_orig:
        raise ValueError("original vector length %d is not a multiple of the %d nodes"
                         % (orig_vec.size, n_orig))
    nv = orig_vec.size // n_orig
    if new_vec is None:
        new_vec = np.zeros(n_new * nv)
    if new_vec.size != n_new * nv or new_vec.dtype != np.double:
        raise ValueError("new vector must hold %d float64 values per node" % nv)
    check(lib().tmrx_forest_interpolate_vec(new_filter.ptr, orig_filter.ptr, nv,
                                            C.c_void_p(orig_vec.ctypes.data),
                                            C.c_void_p(new_vec.ctypes.data), 0),
          "interpolateDesignVec")
    return new_vec


def densityBasedRefine(forest, density, lower=0.05, upper=0.5, reverse=False, min_lev=0,
                       max_lev=TMR.MAX_LEVEL):
    """TopOptUtils.densityBasedRefine (tmr/TopOptUtils.py:397-457): elements
    whose density reaches ``upper`` are refined and those at or below
    ``lower`` coarsened (reversed with ``reverse``), then forest.refine.
    ``density`` holds one value per element, or (nelem, k) values per element
    (TMR takes the max over an element's design values, the min when
    reversed)."""
    d = np.asarray(density, dtype=np.double)
    if d.ndim == 2:
        value = d.min(axis=1) if reverse else d.max(axis=1)
    else:
        value = d
    refine = np.zeros(len(value), dtype=np.intc)
    if reverse:
        refine[value >= upper] = -1
        refine[value <= lower] = 1
    else:
        refine[value >= upper] = 1
        refine[value <= lower] = -1
    forest.refine(refine, min_lev, max_lev)
    return refine


# ---------------------------------------------------------------------------
# M-filter (tmr/TopOptUtils.py:643-890): Helmholtz partition-of-unity filter
# whose stencils are the least-squares weights that reproduce the second-
# order Taylor moments of r^2 times the Laplacian with non-negative weights.
class OptFilterWeights:
    """Stencil-weight problem of one node (tmr/TopOptUtils.py:643-782):
    minimise 1/2 |w|^2 over w >= 0 (one weight per neighbour) subject to
    A w = b, where the columns of A are the neighbour offsets d and the
    second moments d d^T / 2 (upper triangle, off-diagonal terms d_i d_j),
    normalised by the stencil radius delta, and b asks for a zero first
    moment and the second moments of H.  alpha_j = w_j / delta^2 and the
    diagonal alpha_0 = 1 + sum_j w_j / delta^2."""

    def __init__(self, diag, X, H):
        X = np.asarray(X, dtype=float)
        if X.ndim == 1:
            X = X[:, None]
        self.diag, self.n = int(diag), X.shape[0]
        d = np.delete(X - X[self.diag], self.diag, axis=0)
        self.delta = float(np.sqrt(np.max((d * d).sum(axis=1))))
        d = d / self.delta
        dim = X.shape[1]
        self.dim = dim
        H = np.asarray(H, dtype=float)
        if dim == 1:
            A = np.vstack([d[:, 0], 0.5 * d[:, 0] ** 2])
            b = np.array([0.0, H[0, 0]])
        elif dim == 2:
            A = np.vstack([d[:, 0], d[:, 1], 0.5 * d[:, 0] ** 2, 0.5 * d[:, 1] ** 2,
                           d[:, 0] * d[:, 1]])
            b = np.array([0.0, 0.0, H[0, 0], H[1, 1], 2.0 * H[0, 1]])
        else:
            A = np.vstack([d[:, 0], d[:, 1], d[:, 2],
                           0.5 * d[:, 0] ** 2, 0.5 * d[:, 1] ** 2, 0.5 * d[:, 2] ** 2,
                           d[:, 1] * d[:, 2], d[:, 0] * d[:, 2], d[:, 0] * d[:, 1]])
            b = np.array([0.0, 0.0, 0.0, H[0, 0], H[1, 1], H[2, 2],
                          2.0 * H[1, 2], 2.0 * H[0, 2], 2.0 * H[0, 1]])
        self.A, self.b = A, b

    def obj_func(self, w):
        return 0.5 * float(np.dot(w, w))

    def obj_func_der(self, w):
        return w

    def con_func(self, w):
        return self.A @ w - self.b

    def con_func_der(self, w):
        return self.A

    def solve(self):
        from scipy.optimize import minimize
        m = self.n - 1
        res = minimize(self.obj_func, np.ones(m), jac=self.obj_func_der, method="SLSQP",
                       bounds=[(0.0, None)] * m,
                       constraints={"type": "eq", "fun": self.con_func,
                                    "jac": self.con_func_der})
        return res.x

    def set_alphas(self, w, alpha):
        alpha[:] = 0.0
        others = [i for i in range(self.n) if i != self.diag]
        alpha[others] = w / self.delta ** 2
        alpha[self.diag] = 1.0 + float(np.sum(w)) / self.delta ** 2


class Mfilter(topo.HelmholtzPUFilter):
    """TopOptUtils.Mfilter (tmr/TopOptUtils.py:785-890): a Helmholtz
    partition-of-unity filter with radius r whose stencils are the
    OptFilterWeights solutions: H = r^2 I at interior nodes; at boundary
    nodes the stencil is fitted in the tangent plane (coordinates along two
    directions orthogonal to the normal, H = r^2 I in 2D).  Call
    initialize() before use."""

    def __init__(self, N, assemblers, filters, vars_per_node=1, dim=3, r=0.01, device=0):
        super().__init__(N, assemblers, filters, vars_per_node, device)
        self.r, self.dim = float(r), int(dim)

    def getInteriorStencil(self, diag, X, alpha):
        X = np.asarray(X, dtype=float).reshape(-1, 3)
        if self.dim == 2:
            X = X[:, :2]
        opt = OptFilterWeights(diag, X, self.r ** 2 * np.eye(3))
        opt.set_alphas(opt.solve(), alpha)

    def getBoundaryStencil(self, diag, normal, X, alpha):
        X = np.asarray(X, dtype=float).reshape(-1, 3)
        normal = np.asarray(normal, dtype=float)
        if self.dim == 2:
            t = np.array([normal[1], -normal[0]])
            Xt = (X[:, :2] - X[diag, :2]) @ t
        else:
            # two in-plane directions from the coordinate axis least aligned
            # with the normal
            e = np.zeros(3)
            e[np.argmin(np.abs(normal))] = 1.0
            t2 = np.cross(e, normal)
            t1 = np.cross(normal, t2)
            Xt = np.stack([(X - X[diag]) @ t1, (X - X[diag]) @ t2], axis=1)
        opt = OptFilterWeights(diag, Xt, self.r ** 2 * np.eye(2))
        opt.set_alphas(opt.solve(), alpha)
