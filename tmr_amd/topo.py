"""TMR's topology-optimisation surface over the native GPU problem.

Mirrors the Cython classes a TMR user builds a problem from -- same names,
argument meaning and error behaviour -- with the TACS objects they wrap
replaced by the native multigrid-preconditioned PCG (libtmrx.so):

- ``BoundaryConditions``   tmr/TMR.pyx:3877-3930, src/TMRBoundaryConditions.cpp:68-116
- ``MaterialProperties``   TACS ``constitutive.MaterialProperties(rho, E, nu, ys)``
- ``StiffnessProperties``  tmr/TMR.pyx:5041-5130 (defaults q=5, eps=0.3,
  k0=1e-6, RAMP), src/topology/TMROctConstitutive.cpp:26-71
- ``OctConstitutive``      tmr/TMR.pyx:5201-5211
- ``OctTopoCreator``       tmr/TMR.pyx:4238-4261, element densities from a
  filter forest (src/topology/TMR_TACSTopoCreator.cpp:55-371)
- ``OctConformTopoCreator`` tmr/TMR.pyx:4289-4315, nodal design variables on
  the analysis forest (src/topology/TMR_TACSTopoCreator.cpp:714-777)
- ``createMg``             tmr/TMR.pyx:4317-4368, src/TMR_RefinementTools.cpp:37-107
- ``LagrangeFilter``       tmr/TMR.pyx:4776-4815 (design space = filter nodes)
- ``ConformFilter``        tmr/TMR.pyx:4817-4852, src/topology/TMRConformFilter.cpp
- ``TopoProblem``          tmr/TMR.pyx:5371-6005, src/topology/TMRTopoProblem.cpp

Design layout: the design vector lives on the independent nodes of the
finest filter forest.  With ``OctTopoCreator`` element densities are
``rho = W x`` with W the element-centre weights, and every gradient is
returned in design space as ``W^T dF/drho`` (the filter transpose of
TMRTopoProblem.cpp:1282).  With ``OctConformTopoCreator`` the design
variables are the analysis forest's own nodal values, interpolated to every
Gauss point (nodal mode of the native problem), and gradients are per node.

Vectors are numpy arrays (TACS ``Vec`` / ParOpt ``PVec`` in the reference).
"""
import numpy as np

from . import solver

SIMP_PENALTY = solver.SIMP
RAMP_PENALTY = solver.RAMP
NATURAL_ORDER = 0
MULTICOLOR_ORDER = 1


class BoundaryConditions:
    """Named Dirichlet conditions (TMRBoundaryConditions).  ``bc_nums=None``
    constrains every displacement component of the named nodes (TACS
    ``addBCs`` with no variable list), with value zero; ``bc_vals`` (one per
    listed component, tmr/TMR.pyx:3898-3939) prescribes non-zero values, which
    the native solve lifts: K_ff u_f = f_f - K_fc u_c."""

    def __init__(self):
        self._bcs = []

    def getNumBoundaryConditions(self):
        return len(self._bcs)

    def addBoundaryCondition(self, aname, bc_nums=None, bc_vals=None):
        if bc_nums is not None and bc_vals is not None and len(bc_nums) != len(bc_vals):
            raise ValueError("Boundary condition lists must be the same length")
        vals = [0.0, 0.0, 0.0]
        if bc_nums is None or len(bc_nums) == 0:
            mask = 7
        else:
            mask = 0
            for k, n in enumerate(bc_nums):
                n = int(n)
                if n < 0 or n > 2:
                    raise ValueError("boundary condition variable %d out of range [0, 3)" % n)
                mask |= 1 << n
                if bc_vals is not None:
                    vals[n] = float(bc_vals[k])
        self._bcs.append((aname, mask, vals))

    def as_list(self):
        """(name, dof bitmask) pairs, bit i = displacement component i."""
        return [(b[0], b[1]) for b in self._bcs if b[0] is not None]

    def has_values(self):
        return any(v != 0.0 for b in self._bcs for v in b[2])

    def values_vector(self, forest):
        """Prescribed value of every dof of the forest's independent nodes
        (zero where nothing is prescribed; later conditions win)."""
        u = np.zeros(3 * forest.getNodeRange()[1])
        for name, mask, vals in self._bcs:
            if name is None:
                continue
            nodes = np.asarray(forest.getNodesWithName(name))
            nodes = nodes[nodes >= 0]
            for d in range(3):
                if mask & (1 << d):
                    u[3 * nodes + d] = vals[d]
        return u


class MaterialProperties:
    """Isotropic material (TACS MaterialProperties: rho, E, nu, ys)."""

    def __init__(self, rho=1.0, E=1.0, nu=0.3, ys=1.0, **kwargs):
        self.rho, self.E, self.nu, self.ys = float(rho), float(E), float(nu), float(ys)

    def getDensity(self):
        return self.rho


class StiffnessProperties:
    """SIMP/RAMP stiffness parametrisation (TMRStiffnessProperties), with the
    keyword aliases of tmr/TMR.pyx:5047-5110."""

    def __init__(self, props, **kwargs):
        if isinstance(props, (list, tuple)):
            if len(props) != 1:
                raise NotImplementedError("multi-material design is not supported")
            props = props[0]
        self.props = props
        self.nmats = 1
        self.q = float(kwargs.get("stiffness_penalty_value", kwargs.get("q", 5.0)))
        self.eps = float(kwargs.get("stress_relax_value", kwargs.get("eps", 0.3)))
        self.k0 = float(kwargs.get("stiffness_offset", kwargs.get("k0", 1e-6)))
        pt = kwargs.get("penalty_type", "RAMP")
        self.penalty_type = SIMP_PENALTY if pt in ("simp", "SIMP", SIMP_PENALTY) else RAMP_PENALTY
        self.ksWeight = float(kwargs.get("ksWeight", 30.0))
        self.qmass = float(kwargs.get("mass_penalty_value", kwargs.get("qmass", 0.0)))
        if kwargs.get("use_project"):
            raise NotImplementedError("projected densities are not supported")

    def getDesignVarsPerNode(self):
        return 1

    def design_lower_bound(self):
        """TMROctConstitutive::getDesignVarRange (TMROctConstitutive.cpp:188-211)."""
        return 1e-3 if self.penalty_type == SIMP_PENALTY else 0.0


class OctConstitutive:
    """Constitutive object of a topology element (TMROctConstitutive)."""

    def __init__(self, props=None, forest=None):
        if props is None:
            raise ValueError("OctConstitutive needs StiffnessProperties")
        self.props = props
        self.forest = forest


def _props_of(obj):
    """StiffnessProperties behind what a createElement callback returned."""
    seen = 0
    while obj is not None and not isinstance(obj, StiffnessProperties) and seen < 4:
        obj = getattr(obj, "props", None) or getattr(obj, "con", None) or \
            getattr(obj, "constitutive", None)
        seen += 1
    if not isinstance(obj, StiffnessProperties):
        raise TypeError("createElement must return an OctConstitutive, StiffnessProperties or an "
                        "object whose .props/.con leads to one")
    return obj


class Assembler:
    """One mesh level: the forest, its boundary conditions, the stiffness
    parametrisation and the element-centre weights from the filter (what
    TMROctTACSCreator::createTACS, TMR_TACSCreator.cpp:332-461, flattens into
    a TACSAssembler)."""

    def __init__(self, forest, filt, bcs, props, weights=None):
        self.forest, self.filter, self.bcs, self.props = forest, filt, bcs, props
        self.nnodes = forest.getNodeRange()[1]
        self.ndof = 3 * self.nnodes
        ext = forest.getElementExtents()
        self.volumes = np.prod(ext, axis=1)
        self.nelem = len(self.volumes)
        self.nodal = weights is None
        if self.nodal:
            # design variable numbers of element i are its nodes
            # (TMROctConformTACSTopoCreator::createElements)
            self.ndv = self.nnodes
            self._mass_w = self._corner_gather(np.repeat(self.volumes[:, None] / 8.0, 8, axis=1))
        else:
            self.rowp, self.cols, self.w, self.ndv = weights
            self._rows = np.repeat(np.arange(self.nelem), np.diff(self.rowp))
        self.problem = None  # the native problem, once createMg has built one

    def _corner_gather(self, t):
        """sum_e t[e, a] onto the independent nodes (hanging corners pass
        their share to their parents with the hanging weights)."""
        conn = np.asarray(self.forest.getMeshConn()).reshape(-1, 8)
        ptr, dc, dw = self.forest.getDepNodeConn()
        g = np.zeros(self.nnodes)
        ind = conn >= 0
        np.add.at(g, conn[ind], t[ind])
        for e, a in zip(*np.nonzero(~ind)):
            d = -conn[e, a] - 1
            np.add.at(g, dc[ptr[d]:ptr[d + 1]], dw[ptr[d]:ptr[d + 1]] * t[e, a])
        return g

    def mass_weights(self):
        """d(mass)/dx per design variable, density 1: element volumes through
        the element-centre weights, or (nodal) sum_e V_e/8 over the corners at
        each node -- the exact 2x2x2 integral of the trilinear density
        (TMROctConstitutive::evalDensity, src/topology/TMROctConstitutive.cpp:217-248)."""
        if self.nodal:
            return self._mass_w
        return self.weights_transpose(self.volumes)

    def to_design_space(self, g):
        """A native gradient in design space: W^T g for element densities,
        unchanged for nodal designs (already per node)."""
        return np.asarray(g) if self.nodal else self.weights_transpose(g)

    def getNumElements(self):
        return self.nelem

    def getNumOwnedNodes(self):
        return self.nnodes

    def getVarsPerNode(self):
        return 3

    def createVec(self):
        return np.zeros(self.ndof)

    def createDesignVec(self):
        return np.zeros(self.ndv)

    def getDesignVars(self, x=None):
        """Initial design x = 0.95 (TMROctConstitutive.cpp:104-109)."""
        if x is None:
            x = self.createDesignVec()
        x[:] = 0.95
        return x

    def getDesignVarRange(self, lb=None, ub=None):
        lb = self.createDesignVec() if lb is None else lb
        ub = self.createDesignVec() if ub is None else ub
        lb[:] = self.props.design_lower_bound()
        ub[:] = 1.0
        return lb, ub

    def element_densities(self, x):
        """rho_e = sum_j w_ej x_j (element-density assemblers)."""
        if self.nodal:
            raise ValueError("a nodal-design assembler has no element densities")
        x = np.asarray(x, dtype=np.double)
        if x.shape != (self.ndv,):
            raise ValueError("design vector has %s entries, expected %d" % (x.shape, self.ndv))
        rho = np.zeros(self.nelem)
        np.add.at(rho, self._rows, self.w * x[self.cols])
        return rho

    def weights_transpose(self, g_rho):
        """W^T g: element sensitivities to design space."""
        g = np.zeros(self.ndv)
        np.add.at(g, self.cols, self.w * np.asarray(g_rho)[self._rows])
        return g


class OctTopoCreator:
    """Creates a level's Assembler with element densities interpolated from
    the filter forest (TMROctTACSTopoCreator).  A subclass may define
    ``createElement(order, octant, index, weights)``; its return value (an
    OctConstitutive, or an element holding one through ``.props``/``.con``)
    names the StiffnessProperties every element of the level uses.
    Without ``createElement`` the ``props`` keyword is used."""

    def __init__(self, bcs, filt, design_vars_per_node=1, props=None, *args, **kwargs):
        if design_vars_per_node != 1:
            raise NotImplementedError("one design variable per node only")
        self.bcs, self.filt, self.props = bcs, filt, props

    def getFilter(self):
        return self.filt

    def createTACS(self, forest, ordering=NATURAL_ORDER):
        forest.createNodes()
        weights = forest.createElementCenterWeights(self.filt)
        props = self.props
        if hasattr(self, "createElement"):
            octs = forest.getOctants()
            rp, cl, w, _ = weights
            props = _props_of(self.createElement(forest.getMeshOrder(), octs[0],
                                                 list(cl[rp[0]:rp[1]]), list(w[rp[0]:rp[1]])))
        if props is None:
            raise ValueError("no StiffnessProperties: pass props= or define createElement")
        return Assembler(forest, self.filt, self.bcs, props, weights)


class OctConformTopoCreator:
    """Nodal design variables on the analysis forest (TMROctConformTACSTopoCreator,
    src/topology/TMR_TACSTopoCreator.cpp:714-777): element i's design
    variables are its 8 nodes; the density at each Gauss point is their
    trilinear interpolation and the penalty applies per point.  The filter is
    the forest itself (``order < 0`` or the forest's own order); another
    order is not built.  A subclass may define
    ``createElement(order, octant, index, filter)`` returning what names the
    StiffnessProperties, as with OctTopoCreator."""

    def __init__(self, bcs, forest, design_vars_per_node=1, order=-1, interp=None, props=None,
                 *args, **kwargs):
        if design_vars_per_node != 1:
            raise NotImplementedError("one design variable per node only")
        if order >= 0 and order != forest.getMeshOrder():
            raise NotImplementedError("a filter order other than the forest's is not supported")
        self.bcs, self.forest, self.props = bcs, forest, props
        forest.createNodes()

    def getFilter(self):
        return self.forest

    def createTACS(self, forest, ordering=NATURAL_ORDER):
        forest.createNodes()
        props = self.props
        if hasattr(self, "createElement"):
            octs = forest.getOctants()
            conn = np.asarray(forest.getMeshConn()).reshape(-1, 8)
            props = _props_of(self.createElement(forest.getMeshOrder(), octs[0],
                                                 list(conn[0]), forest))
        if props is None:
            raise ValueError("no StiffnessProperties: pass props= or define createElement")
        return Assembler(forest, forest, self.bcs, props, None)


class Mg:
    """The multigrid-preconditioned solver of a level hierarchy (TACSMg from
    TMR_CreateTACSMg, src/TMR_RefinementTools.cpp:37-107).  ``use_galerkin``
    forms the coarse operators P^T A P on the GPU (tmrx_mg_create).  The
    smoother is TMR's default symmetric SOR with relaxation ``omega`` on
    the assembled level matrices (``use_chebyshev_smoother=False``, swept in
    colour-major node order), or Chebyshev on the Jacobi-scaled operator;
    the coarsest level is a direct FP64 Cholesky solve, or with
    ``use_coarse_direct_solve=False`` the level's smoother once (TMR hands
    TACSMg the smoother there, src/TMR_RefinementTools.cpp:87-102)."""

    def __init__(self, assemblers, forests, omega=1.0, use_galerkin=False,
                 use_coarse_direct_solve=True, use_chebyshev_smoother=False, device=0):
        if len(assemblers) != len(forests) or len(assemblers) < 1:
            raise ValueError("createMg: list length mismatch")
        a0 = assemblers[0]
        p, m = a0.props, a0.props.props
        self.assemblers, self.forests = list(assemblers), list(forests)
        self.problem = solver.Problem(self.forests, E=m.E, nu=m.nu, ys=m.ys, q=p.q, k0=p.k0,
                                      eps=p.eps, penalty_type=p.penalty_type,
                                      bcs=a0.bcs.as_list(), device=device)
        self.nlevels = len(assemblers)
        a0.problem = self.problem
        if a0.bcs.has_values():
            self.problem.set_bc_values(a0.bcs.values_vector(self.forests[0]))
        self.problem.set_multigrid(omega=omega, use_galerkin=use_galerkin,
                                   use_coarse_direct_solve=use_coarse_direct_solve,
                                   use_chebyshev_smoother=use_chebyshev_smoother)
        self.opts = solver.default_opts()

    def getAssembler(self):
        return self.assemblers[0]


def createMg(assemblers, forests, omega=1.0, use_galerkin=False, use_coarse_direct_solve=True,
             use_chebyshev_smoother=False):
    return Mg(assemblers, forests, omega, use_galerkin, use_coarse_direct_solve,
              use_chebyshev_smoother)


class TopoFilter:
    def __init__(self, assemblers, filters):
        self.assemblers, self.filters = list(assemblers), list(filters)
        self.x = None

    def getFilter(self):
        return self.filters[0]

    def getAssembler(self):
        return self.assemblers[0]

    def setDesignVars(self, x):
        self.x = np.array(x, dtype=np.double)

    def getDesignVars(self, x=None):
        if x is None:
            return None if self.x is None else self.x.copy()
        x[:] = self.x
        return x

    def applyFilter(self, vec_in, vec_out=None):
        """Element densities of the finest level from a design vector (the
        design vector itself for a nodal design)."""
        a = self.assemblers[0]
        rho = np.array(vec_in, dtype=np.double) if a.nodal else a.element_densities(vec_in)
        if vec_out is not None:
            vec_out[:] = rho
            return vec_out
        return rho

    def addValues(self, g_rho):
        """Filter transpose of an element sensitivity (identity for a nodal
        design)."""
        return self.assemblers[0].to_design_space(g_rho)

    def filter_nodal(self, x):
        """The nodal design the analysis sees (TMRConformFilter: x itself)."""
        return x

    def transpose_nodal(self, g):
        """Adjoint of filter_nodal for a per-node gradient."""
        return g


class LagrangeFilter(TopoFilter):
    """Design variables on the finest filter forest's nodes
    (TMRLagrangeFilter with OctTopoCreator levels)."""


class ConformFilter(TopoFilter):
    """Nodal design variables on the analysis forests (TMRConformFilter,
    src/topology/TMRConformFilter.cpp): the design vector of level k+1 is the
    weighted restriction (P^T x_k) / (P^T 1) of level k's, done natively by
    the problem when the design is set."""

    def __init__(self, assemblers, filters):
        if len(assemblers) != len(filters):
            raise ValueError("ConformFilter must have equal number of objects in lists")
        if not all(getattr(a, "nodal", False) for a in assemblers):
            raise ValueError("ConformFilter needs OctConformTopoCreator assemblers")
        super().__init__(assemblers, filters)


class _NativeFilter(ConformFilter):
    """A TMRConformFilter whose analysis design is the filtered design
    variables: the filter and its transpose run natively on the finest
    filter forest (tmrx_filter_*)."""

    def __init__(self, kind, r, nterms, assemblers, filters, vars_per_node=1, device=0):
        import ctypes as C
        from ._lib import check, lib
        super().__init__(assemblers, filters)
        if vars_per_node != 1:
            raise NotImplementedError("one design variable per node only")
        self.vars_per_node = vars_per_node
        h = C.c_void_p()
        check(lib().tmrx_filter_create(self.filters[0].ptr, int(kind), float(r), int(nterms),
                                       int(device), C.byref(h)), type(self).__name__)
        self._h = h
        n = C.c_int()
        check(lib().tmrx_filter_info(self._h, C.byref(n), None))
        self.num_nodes = n.value

    def __del__(self):
        try:
            from ._lib import lib
            if getattr(self, "_h", None):
                lib().tmrx_filter_free(self._h)
                self._h = None
        except Exception:
            pass

    def _apply(self, x, transpose):
        import ctypes as C
        from ._lib import check, lib
        x = np.ascontiguousarray(x, dtype=np.double)
        if x.size != self.num_nodes * self.vars_per_node:
            raise ValueError("expected %d values, got %d" % (self.num_nodes * self.vars_per_node,
                                                               x.size))
        out = np.zeros_like(x)
        fn = lib().tmrx_filter_apply_transpose if transpose else lib().tmrx_filter_apply
        check(fn(self._h, self.vars_per_node, C.c_void_p(x.ctypes.data),
                 C.c_void_p(out.ctypes.data), 0), "filter")
        return out

    def applyFilter(self, vec_in, vec_out=None):
        """TMRTopoFilter::applyFilter (src/topology/TMRTopoFilter.h:59-62)."""
        out = self._apply(vec_in, False)
        if vec_out is not None:
            vec_out[:] = out
            return vec_out
        return out

    def applyTranspose(self, vec_in, vec_out=None):
        """TMRTopoFilter::applyTranspose (src/topology/TMRTopoFilter.h:63-66)."""
        out = self._apply(vec_in, True)
        if vec_out is not None:
            vec_out[:] = out
            return vec_out
        return out

    def filter_nodal(self, x):
        return self._apply(x, False)

    def transpose_nodal(self, g):
        return self._apply(g, True)

    def last_iterations(self):
        import ctypes as C
        from ._lib import check, lib
        it = C.c_int()
        check(lib().tmrx_filter_info(self._h, None, C.byref(it)))
        return it.value


class HelmholtzFilter(_NativeFilter):
    """TMRHelmholtzFilter (src/topology/TMRHelmholtzFilter.cpp; Python
    TMR.HelmholtzFilter(r, assemblers, filters, vars_per_node)): the
    analysis design is (M + r^2 K)^-1 M x on the finest filter forest."""

    def __init__(self, r, assemblers, filters, vars_per_node=1, device=0):
        super().__init__(0, r, 0, assemblers, filters, vars_per_node, device)


def _filter_pattern(forest):
    """Rows of the filter forest's node adjacency: for every independent node
    the sorted independent nodes that share an element with it, hanging
    corners replaced by their parents (the pattern of the assembled
    matrices TMRHelmholtzPUFilter::initialize stencils over,
    src/topology/TMRHelmholtzPUFilter.cpp:437-512)."""
    conn = np.asarray(forest.getMeshConn()).reshape(-1, 8)
    dp, dc, dw = forest.getDepNodeConn()
    n = int(forest.getNodeRange()[1])
    # the parents of every corner, padded with -1 (at most 4)
    par = np.full((conn.shape[0], 8, 4), -1, dtype=np.int64)
    ind = conn >= 0
    par[..., 0] = np.where(ind, conn, -1)
    d = np.where(ind, 0, -conn - 1)
    for k in range(4):
        take = (~ind) & (dp[d] + k < dp[d + 1])
        par[..., k] = np.where(take, dc[np.minimum(dp[d] + k, len(dc) - 1)], par[..., k])
    nodes = par.reshape(conn.shape[0], 32)
    a = np.repeat(nodes, 32, axis=1).ravel()
    b = np.tile(nodes, (1, 32)).ravel()
    keep = (a >= 0) & (b >= 0)
    key = np.unique(a[keep] * n + b[keep])
    rows, cols = key // n, key % n
    rowp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowp, rows + 1, 1)
    return np.cumsum(rowp).astype(np.intc), cols.astype(np.intc)


def _boundary_normals(forest, P):
    """Sum over the element faces on the domain boundary of their outward
    unit normals at each independent corner node on them (zero inside):
    computeOctreeBoundaryNormals (src/topology/TMRHelmholtzPUFilter.cpp:
    183-302), for elements that are parallelepipeds (box blocks)."""
    octs = forest.getOctants()
    conn = np.asarray(forest.getMeshConn()).reshape(-1, 8)
    nums, X = forest.getNodeNumbers(), forest.getPoints()
    ndep = int((nums < 0).sum())
    cinfo = forest.getConnectivity()
    bfc = np.asarray(cinfo["block_face_conn"]).reshape(-1, 6)
    nblk_face = np.bincount(bfc.ravel(), minlength=int(cinfo["nfaces"]))  # blocks per face
    hmax = 1 << 30
    h = np.left_shift(1, 30 - octs["level"].astype(np.int64))
    coord = np.stack([octs["x"], octs["y"], octs["z"]], 1).astype(np.int64)
    N = np.zeros((len(P), 3))
    for e in range(len(octs)):
        b = octs["block"][e]
        Xe = X[conn[e] + ndep]  # corner points (getPoints order: dependent first)
        for surf in range(6):
            ax, hi = surf // 2, surf % 2
            if nblk_face[bfc[b, surf]] != 1:
                continue
            if coord[e, ax] != (hmax - h[e] if hi else 0):
                continue
            # local tangents of the face (d1, d2 as the reference orders them)
            t1, t2 = [(1, 2), (2, 0), (0, 1)][ax]
            da = Xe[1 << t1] - Xe[0]
            db = Xe[1 << t2] - Xe[0]
            nrm = np.cross(da, db)
            nrm /= np.linalg.norm(nrm)
            if hi == 0:
                nrm = -nrm
            for c in range(8):
                if ((c >> ax) & 1) == hi and conn[e, c] >= 0:
                    N[conn[e, c]] += nrm
    return N


class HelmholtzPUFilter(_NativeFilter):
    """TMRHelmholtzPUFilter (src/topology/TMRHelmholtzPUFilter.cpp; Python
    TMR.HelmholtzPUFilter(N, assemblers, filters, vars_per_node)): a
    partition-of-unity filter from a stencil at every node of the finest
    filter forest.  Subclasses provide getInteriorStencil(diag, X, alpha) and
    getBoundaryStencil(diag, normal, X, alpha) (X: the stencil points as 3n
    values, alpha: the n weights to fill); initialize() calls them row by row
    and builds F = T^-1 sum_{k=0}^{N} (D^-1 B)^k D^-1 natively (D the
    diagonal weights, B the others clipped at zero)."""

    def __init__(self, N, assemblers, filters, vars_per_node=1, device=0):
        ConformFilter.__init__(self, assemblers, filters)
        if vars_per_node != 1:
            raise NotImplementedError("one design variable per node only")
        self.vars_per_node = vars_per_node
        self.N, self.device = int(N), int(device)
        self._h = None
        self.num_nodes = int(self.filters[0].getNodeRange()[1])

    def getInteriorStencil(self, diag, X, alpha):
        raise NotImplementedError("HelmholtzPUFilter subclasses provide the stencils")

    def getBoundaryStencil(self, diag, normal, X, alpha):
        raise NotImplementedError("HelmholtzPUFilter subclasses provide the stencils")

    def stencils(self):
        """(rowp, cols, B, D) of the filter from the subclass's stencils."""
        f = self.filters[0]
        rowp, cols = _filter_pattern(f)
        nums, X = f.getNodeNumbers(), f.getPoints()
        P = np.zeros((self.num_nodes, 3))
        P[nums[nums >= 0]] = X[nums >= 0]
        normals = _boundary_normals(f, P)
        B = np.zeros(len(cols))
        D = np.ones(self.num_nodes)
        for i in range(self.num_nodes):
            c = cols[rowp[i]:rowp[i + 1]]
            diag = int(np.nonzero(c == i)[0][0])
            Xs = P[c].reshape(-1).copy()
            alpha = np.zeros(len(c))
            nrm = normals[i]
            if not nrm.any():
                self.getInteriorStencil(diag, Xs, alpha)
            else:
                self.getBoundaryStencil(diag, nrm / np.linalg.norm(nrm), Xs, alpha)
            D[i] = alpha[diag] if alpha[diag] > 0.0 else 1.0
            w = np.maximum(alpha, 0.0)
            w[diag] = 0.0
            B[rowp[i]:rowp[i + 1]] = w
        return rowp, cols, B, D

    def initialize(self):
        """TMRHelmholtzPUFilter::initialize: the stencils, then the native
        filter (and its normalisation T^-1 = 1 / F(1))."""
        import ctypes as C
        from ._lib import check, lib
        rowp, cols, B, D = self.stencils()
        self._stencil = (rowp, cols, B, D)
        h = C.c_void_p()
        check(lib().tmrx_filter_create_custom(self.num_nodes, self.N, rowp.ctypes.data,
                                              cols.ctypes.data, B.ctypes.data, D.ctypes.data,
                                              self.device, C.byref(h)), type(self).__name__)
        self._h = h

    def _apply(self, x, transpose):
        if self._h is None:
            raise RuntimeError("HelmholtzPUFilter: call initialize() first")
        return _NativeFilter._apply(self, x, transpose)


class MatrixFilter(_NativeFilter):
    """TMRMatrixFilter (src/topology/TMRMatrixFilter.cpp; Python
    TMR.MatrixFilter(r, N, assemblers, filters, vars_per_node)): the N-term
    polynomial mass-matrix filter."""

    def __init__(self, r, N, assemblers, filters, vars_per_node=1, device=0):
        super().__init__(1, r, N, assemblers, filters, vars_per_node, device)


# ---------------------------------------------------------------------------
# functions of the state (TACS Compliance / StructuralMass / stress failure)
class Compliance:
    """u . f of a load case."""


class StructuralMass:
    """Integral of rho times the material density (TMROctConstitutive::evalDensity,
    TMROctConstitutive.cpp:217-279): sum_e V_e rho_e for element densities,
    the exact integral of the trilinear nodal density otherwise."""


class PNormFailure:
    """Relaxed von Mises p-norm (f = (sum_e sum_g V_e/8 fail^P)^(1/P),
    fail = rho/(eps(1-rho)+rho) vm(C0 B u)/ys)."""

    def __init__(self, P=8.0):
        self.P = float(P)


class TopoProblem:
    """The optimisation problem (TMRTopoProblem): load cases, a weighted
    compliance or function objective, constraints ``scale*(f + offset)``,
    linear constraints and callbacks, evaluated through the native solver."""

    def __init__(self, fltr, pc=None, gmres_subspace=50, rtol=1e-9):
        if pc is None or not isinstance(pc, Mg):
            raise ValueError("TopoProblem requires a Mg preconditioner")
        self.filter, self.mg = fltr, pc
        self.assembler = fltr.getAssembler()
        self.problem = pc.problem
        self.rtol = float(rtol)
        self.forces, self.vars = [], []
        self.obj_weights, self.obj_funcs = None, None
        self.cons = []  # (case, funcs, offset, scale)
        self.lin_vecs, self.lin_offset = [], []
        self.objcb = self.objgradcb = None
        self.concb = []  # (ncon, confunc, congradfunc)
        self.xinit = self.xlb = self.xub = None
        self.prefix, self.iter_count = "", 0
        self.initialized = False
        self.last_info = []

    # --- configuration ----------------------------------------------------
    def getAssembler(self):
        return self.assembler

    def getFilter(self):
        return self.filter.getFilter()

    def getTopoFilter(self):
        return self.filter

    def getMg(self):
        return self.mg

    def setLoadCases(self, forces):
        self.forces = [None if f is None else np.array(f, dtype=np.double) for f in forces]
        for f in self.forces:
            if f is not None and f.shape != (self.assembler.ndof,):
                raise ValueError("load vector has %s entries, expected %d" % (f.shape,
                                                                           self.assembler.ndof))
        self.vars = [np.zeros(self.assembler.ndof) for _ in forces]
        if self.obj_weights is None:
            self.obj_weights = np.ones(len(forces))

    def getNumLoadCases(self):
        return len(self.forces)

    def setObjective(self, weights, funcs=None):
        if len(weights) != self.getNumLoadCases():
            raise ValueError("Incorrect number of weights")
        self.obj_weights = np.array(weights, dtype=np.double)
        self.obj_funcs = list(funcs) if funcs else None

    def addConstraints(self, case, funcs, offset, scale):
        if case < 0 or case >= self.getNumLoadCases():
            raise ValueError("Load case out of expected range")
        if len(funcs) != len(offset) or len(funcs) != len(scale):
            raise ValueError("Expected equal function, offset and scale counts")
        self.cons.append((case, list(funcs), [float(o) for o in offset], [float(s) for s in scale]))

    def addLinearConstraints(self, vecs, offset):
        if len(vecs) != len(offset):
            raise ValueError("Expected equal vector and offset counts")
        self.lin_vecs += [np.array(v, dtype=np.double) for v in vecs]
        self.lin_offset += [float(o) for o in offset]

    def addObjectiveCallback(self, objfunc, objgradfunc):
        self.objcb, self.objgradcb = objfunc, objgradfunc

    def addConstraintCallback(self, ncon, nineq, confunc, congradfunc):
        self.concb.append((int(ncon), confunc, congradfunc))

    def setPrefix(self, prefix):
        self.prefix = prefix

    def setIterationCounter(self, count):
        self.iter_count = int(count)

    def setInitDesignVars(self, x, lb=None, ub=None):
        self.xinit = np.array(x, dtype=np.double)
        self.xlb = None if lb is None else np.array(lb, dtype=np.double)
        self.xub = None if ub is None else np.array(ub, dtype=np.double)

    def initialize(self):
        self.initialized = True

    def getNumConstraints(self):
        n = len(self.lin_vecs) + sum(len(c[1]) for c in self.cons)
        return n + sum(c[0] for c in self.concb)

    # --- ParOpt problem interface --------------------------------------------
    def createDesignVec(self):
        return self.assembler.createDesignVec()

    def getVarsAndBounds(self, x=None, lb=None, ub=None):
        """TMRTopoProblem::getVarsAndBounds (TMRTopoProblem.cpp:911-950)."""
        a = self.assembler
        x = a.createDesignVec() if x is None else x
        if self.xinit is not None:
            x[:] = self.xinit
        else:
            a.getDesignVars(x)
        lb0, ub0 = a.getDesignVarRange()
        lb = lb0 if lb is None else lb
        ub = ub0 if ub is None else ub
        lb[:] = self.xlb if self.xlb is not None else lb0
        ub[:] = self.xub if self.xub is not None else ub0
        return x, lb, ub

    def _to_design(self, g_native):
        """Native gradient (per element or per analysis node) to the design
        space: the creator's weights, then the filter transpose."""
        g = self.assembler.to_design_space(g_native)
        return self.filter.transpose_nodal(g) if self.assembler.nodal else g

    def _eval_func(self, func, case):
        p = self.problem
        if isinstance(func, Compliance):
            return float(np.dot(self.vars[case], self.forces[case]))
        if isinstance(func, StructuralMass):
            a = self.assembler
            if a.nodal:
                return float(np.dot(a.mass_weights(), self.x_design)) * a.props.props.getDensity()
            return float(np.dot(a.volumes, self.rho)) * a.props.props.getDensity()
        if isinstance(func, PNormFailure):
            self._restore(case)
            return p.stress_pnorm(func.P)
        raise TypeError("unsupported function %r" % (func,))

    def _eval_func_grad(self, func, case):
        """Native gradient of func for load case `case`: per element
        (element densities) or per node (nodal design)."""
        p = self.problem
        if isinstance(func, Compliance):
            self._restore(case)
            return p.compliance_grad()
        if isinstance(func, StructuralMass):
            a = self.assembler
            d = a.props.props.getDensity()
            return a.mass_weights() * d if a.nodal else a.volumes * d
        if isinstance(func, PNormFailure):
            self._restore(case)
            return p.stress_pnorm_grad(func.P, opts=self._opts())
        raise TypeError("unsupported function %r" % (func,))

    def _opts(self):
        return solver.default_opts(rtol=self.rtol, precision=1)

    def _restore(self, case):
        """Make load case `case` the native problem's current state."""
        if self._current == case:
            return
        self.problem.set_state(self.forces[case], self.vars[case])
        self._current = case

    def setDesignVars(self, x):
        """TMRTopoProblem::setDesignVars (TMRTopoProblem.cpp:955-962)."""
        self.filter.setDesignVars(x)
        self.x_design = np.array(x, dtype=np.double)
        if self.assembler.nodal:
            self.rho = None
            # the analysis design (filtered by a Helmholtz or matrix filter)
            self.x_design = np.array(self.filter.filter_nodal(self.x_design), dtype=np.double)
            self.problem.set_design_nodal(self.x_design)
        else:
            self.rho = self.assembler.element_densities(x)
            self.problem.set_design(self.rho)
        self._current = None

    def evalObjCon(self, ncon, x):
        """TMRTopoProblem::evalObjCon (TMRTopoProblem.cpp:967-1060): returns
        (fail, fobj, cons)."""
        self.setDesignVars(x)
        fobj = 0.0
        cons = np.zeros(ncon)
        k = 0
        for v, o in zip(self.lin_vecs, self.lin_offset):
            cons[k] = o + float(np.dot(v, x))
            k += 1
        if self.objcb is not None:
            fobj = float(self.objcb(self.filter, self.mg))
        fail = 0
        self.last_info = []
        if self.forces:
            self.problem.setup(self._opts())
            # every load case in one batched solve (TMRTopoProblem.cpp:1003-1009
            # solves them one after another)
            cases = [i for i, f in enumerate(self.forces) if f is not None]
            if cases:
                U = self.problem.solve_multi(np.stack([self.forces[i] for i in cases]),
                                             opts=self._opts(), allow_unconverged=True)
                for j, i in enumerate(cases):
                    self.vars[i] = np.array(U[j])
                    info = self.problem.infos[j]
                    self.last_info.append((info.iters, info.converged))
                    if not info.converged:
                        fail = 1
                self._current = cases[0]
            for i, f in enumerate(self.forces):
                if f is None:
                    continue
                if self.obj_funcs and self.obj_funcs[i] is not None:
                    fobj += self.obj_weights[i] * self._eval_func(self.obj_funcs[i], i)
                elif not self.obj_funcs:
                    fobj += self.obj_weights[i] * float(np.dot(self.vars[i], f))
                for case, funcs, offs, scales in self.cons:
                    if case != i:
                        continue
                    for fn, o, s in zip(funcs, offs, scales):
                        cons[k] = s * (self._eval_func(fn, i) + o)
                        k += 1
        for ncb, cf, _ in self.concb:
            vals = cf(self.filter, self.mg)
            cons[k:k + ncb] = vals
            k += ncb
        self.x = np.array(x, dtype=np.double)
        return fail, fobj, cons

    def evalObjConGradient(self, x, g, A):
        """TMRTopoProblem::evalObjConGradient (TMRTopoProblem.cpp:1216-1295):
        fills the design-space gradient g and the constraint gradients A."""
        a = self.assembler
        g[:] = 0.0
        if self.objgradcb is not None:
            self.objgradcb(self.filter, self.mg, g)
        grho = np.zeros(a.ndv if a.nodal else a.nelem)
        for i, f in enumerate(self.forces):
            if f is None or self.obj_weights[i] == 0.0:
                continue
            fn = self.obj_funcs[i] if self.obj_funcs else Compliance()
            if fn is None:
                continue
            grho += self.obj_weights[i] * self._eval_func_grad(fn, i)
        g += self._to_design(grho)
        k = 0
        for v in self.lin_vecs:
            A[k][:] = v
            k += 1
        for case, funcs, offs, scales in self.cons:
            for fn, s in zip(funcs, scales):
                A[k][:] = s * self._to_design(self._eval_func_grad(fn, case))
                k += 1
        for ncb, _, cgf in self.concb:
            cgf(self.filter, self.mg, A[k:k + ncb])
            k += ncb
        return 0
