"""TMR's TopOptUtils helpers (tmr/TopOptUtils.py:14-267) over the native
problem: the level-hierarchy builder and the load vectors.

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

    Only order-2 meshes and the "lagrange" (element densities interpolated
    from the filter forest, OctTopoCreator) and "conform" (nodal design
    variables on the analysis forests, OctConformTopoCreator) filters are
    built."""
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


def oc_update(x, g, gcon, vol_frac_excess, lb, ub, move=0.2, eta=0.5):
    """One optimality-criteria step for min compliance s.t. a volume bound,
    with bisection on the multiplier: x_new = clip(x (-g / (lam gcon))^eta).
    A plain driver for examples; the reference hands the problem to ParOpt.
    ``vol_frac_excess(x)`` is the constraint in the form c(x) <= 0."""
    gneg = np.maximum(-g, 1e-30)
    lo_l, hi_l = 1e-12, 1e12
    for _ in range(200):
        lam = np.sqrt(lo_l * hi_l)
        xn = np.clip(x * (gneg / (lam * np.maximum(gcon, 1e-30))) ** eta,
                     np.maximum(lb, x - move), np.minimum(ub, x + move))
        if vol_frac_excess(xn) > 0:
            lo_l = lam
        else:
            hi_l = lam
        if hi_l / lo_l < 1 + 1e-10:
            break
    return xn
