"""ctypes front end of oracle/liboracle.so (test infrastructure only)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build
            build.build_oracle()
        L = C.CDLL(path)
        D, I, P = C.c_double, C.c_int, C.c_void_p
        L.orc_hex8_stiffness.argtypes = [D, D, D, D, D, P]
        L.orc_penalty.restype = D
        L.orc_penalty.argtypes = [D, D, D, I]
        L.orc_dpenalty.restype = D
        L.orc_dpenalty.argtypes = [D, D, I]
        L.orc_assemble.restype = C.c_long
        L.orc_assemble.argtypes = [I, I, P, P, P, P, P, P, P, D, D]
        L.orc_get_csr.argtypes = [P, P, P]
        L.orc_matvec.argtypes = [P, P]
        L.orc_solve.restype = I
        L.orc_solve.argtypes = [P, P, D, I, P]
        L.orc_elem_energy.argtypes = [I, P, P, P, P, P, P, D, D, P, P]
        L.orc_elem_energy2.argtypes = [I, P, P, P, P, P, P, D, D, P, P, P]
        L.orc_hex8_stiffness_gp.argtypes = [D, D, D, D, D, I, P]
        L.orc_set_threads.argtypes = [I]
        L.orc_assemble_gp.restype = C.c_long
        L.orc_assemble_gp.argtypes = [I, I, P, P, P, P, P, P, P, D, D]
        L.orc_elem_energy_gp.argtypes = [I, P, P, P, P, P, P, D, D, P, P, P]
        L.orc_stress_pnorm_gp.restype = D
        L.orc_stress_pnorm_gp.argtypes = [I, I, P, P, P, P, P, P, D, D, D, D, D, P, P, P, P]
        _L = L
    return _L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def hex8_stiffness(hx, hy, hz, E=1.0, nu=0.3):
    K = np.zeros((24, 24))
    lib().orc_hex8_stiffness(hx, hy, hz, E, nu, _p(K))
    return K


def set_threads(n):
    """OpenMP threads of the oracle's CG (1 by default)."""
    lib().orc_set_threads(int(n))


def hex8_stiffness_gp(hx, hy, hz, q, E=1.0, nu=0.3):
    """Contribution of Gauss point q (bit d set: coordinate d at +1/sqrt(3))."""
    K = np.zeros((24, 24))
    lib().orc_hex8_stiffness_gp(hx, hy, hz, E, nu, int(q), _p(K))
    return K


def penalty(rho, q, k0, ptype=0):
    return lib().orc_penalty(rho, q, k0, ptype)


def dpenalty(rho, q, ptype=0):
    return lib().orc_dpenalty(rho, q, ptype)


# the Oracle whose matrix the C library holds (it keeps one assembled
# operator at a time)
_HOLDER = [None]


class Oracle:
    """Assembled FP64 operator A = Z^T K(rho) Z with clamped dofs.  The C
    library holds one assembled operator; an Oracle whose matrix another
    one has replaced assembles its own again before reading it."""

    def __init__(self, conn, dep_ptr, dep_conn, dep_w, hdim, bcmask, E=1.0, nu=0.3):
        self.conn = np.ascontiguousarray(conn, dtype=np.intc).reshape(-1)
        self.dep_ptr = np.ascontiguousarray(dep_ptr, dtype=np.intc)
        if self.dep_ptr.size == 0:
            self.dep_ptr = np.zeros(1, np.intc)
        self.dep_conn = np.ascontiguousarray(dep_conn, dtype=np.intc)
        if self.dep_conn.size == 0:
            self.dep_conn = np.zeros(1, np.intc)
        self.dep_w = np.ascontiguousarray(dep_w, dtype=np.double)
        if self.dep_w.size == 0:
            self.dep_w = np.zeros(1)
        self.hdim = np.ascontiguousarray(hdim, dtype=np.double).reshape(-1)
        self.bcmask = np.ascontiguousarray(bcmask, dtype=np.uint8)
        self.nnodes = len(self.bcmask)
        self.nelem = len(self.conn) // 8
        self.E, self.nu = E, nu

    def assemble(self, s):
        s = np.ascontiguousarray(s, dtype=np.double)
        nnz = lib().orc_assemble(self.nnodes, self.nelem, _p(self.conn), _p(self.dep_ptr),
                                 _p(self.dep_conn), _p(self.dep_w), _p(self.hdim), _p(s),
                                 _p(self.bcmask), self.E, self.nu)
        self.nnz = nnz
        self._made = (self.assemble, s)
        _HOLDER[0] = self
        return nnz

    def _hold(self):
        """Make this Oracle's matrix the library's again if another Oracle
        has assembled since."""
        if _HOLDER[0] is not self:
            if not hasattr(self, "_made"):
                raise RuntimeError("oracle: assemble first")
            fn, arg = self._made
            fn(arg)

    def csr(self):
        self._hold()
        n = 3 * self.nnodes
        rowp = np.zeros(n + 1, np.intc)
        cols = np.zeros(self.nnz, np.intc)
        vals = np.zeros(self.nnz)
        lib().orc_get_csr(_p(rowp), _p(cols), _p(vals))
        return rowp, cols, vals

    def matvec(self, x):
        self._hold()
        x = np.ascontiguousarray(x, dtype=np.double)
        y = np.zeros_like(x)
        lib().orc_matvec(_p(x), _p(y))
        return y

    def solve(self, b, rtol=1e-12, maxit=20000):
        self._hold()
        b = np.ascontiguousarray(b, dtype=np.double).copy()
        b[np.repeat(self.bcmask, 3).reshape(-1, 3).astype(bool).reshape(-1) &
          (((self.bcmask[:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))] = 0.0
        x = np.zeros_like(b)
        rr = C.c_double()
        it = lib().orc_solve(_p(b), _p(x), rtol, maxit, C.byref(rr))
        self.iters, self.relres = it, rr.value
        if it < 0:
            raise RuntimeError("oracle CG did not converge")
        return x

    def apply_free(self, s, x):
        """y = A x element by element (no assembly): A = Z^T (sum_e s_e K_e) Z
        with identity rows and columns on clamped dofs."""
        L = lib()
        if not hasattr(L, "_free_sigs"):
            P, I, D = C.c_void_p, C.c_int, C.c_double
            L.orc_apply_free.argtypes = [I, I, P, P, P, P, P, P, P, D, D, P, P]
            L._free_sigs = True
        s = np.ascontiguousarray(s, dtype=np.double)
        x = np.ascontiguousarray(x, dtype=np.double)
        y = np.zeros(3 * self.nnodes)
        L.orc_apply_free(self.nnodes, self.nelem, _p(self.conn), _p(self.dep_ptr),
                         _p(self.dep_conn), _p(self.dep_w), _p(self.hdim), _p(s),
                         _p(self.bcmask), self.E, self.nu, _p(x), _p(y))
        return y

    def free_mask(self):
        """True on the free dofs."""
        return ~(((self.bcmask[:, None] >> np.arange(3)) & 1).astype(bool).reshape(-1))

    def true_residual(self, s, f, u):
        """|f - A u| / |f| over the free dofs, f zeroed on clamped dofs (the
        oracle's check of a solution, any size)."""
        fr = np.where(self.free_mask(), f, 0.0)
        r = fr - self.apply_free(s, u)
        return float(np.linalg.norm(r) / np.linalg.norm(fr))

    def energy2(self, u, v):
        u = np.ascontiguousarray(u, dtype=np.double)
        v = np.ascontiguousarray(v, dtype=np.double)
        e = np.zeros(self.nelem)
        lib().orc_elem_energy2(self.nelem, _p(self.conn), _p(self.dep_ptr), _p(self.dep_conn),
                               _p(self.dep_w), _p(self.hdim), _p(self.bcmask), self.E, self.nu,
                               _p(u), _p(v), _p(e))
        return e

    def assemble_gp(self, sgp):
        """Nodal design: K_e = sum_q sgp[e, q] K_q (src/topology/TMROctConstitutive.cpp:498-551)."""
        sgp = np.ascontiguousarray(sgp, dtype=np.double).reshape(-1)
        if sgp.size != 8 * self.nelem:
            raise ValueError("need 8 scales per element")
        nnz = lib().orc_assemble_gp(self.nnodes, self.nelem, _p(self.conn), _p(self.dep_ptr),
                                    _p(self.dep_conn), _p(self.dep_w), _p(self.hdim), _p(sgp),
                                    _p(self.bcmask), self.E, self.nu)
        self.nnz = nnz
        self._made = (self.assemble_gp, sgp)
        _HOLDER[0] = self
        return nnz

    def energy_gp(self, u, v):
        """v_e^T K_q u_e per element and Gauss point, shape (nelem, 8)."""
        u = np.ascontiguousarray(u, dtype=np.double)
        v = np.ascontiguousarray(v, dtype=np.double)
        e = np.zeros(8 * self.nelem)
        lib().orc_elem_energy_gp(self.nelem, _p(self.conn), _p(self.dep_ptr), _p(self.dep_conn),
                                 _p(self.dep_w), _p(self.hdim), _p(self.bcmask), self.E, self.nu,
                                 _p(u), _p(v), _p(e))
        return e.reshape(-1, 8)

    def corner_values(self, x):
        """Nodal design values at every element corner (hanging corners
        interpolated from their parents), shape (nelem, 8)."""
        x = np.asarray(x, dtype=np.double)
        c = self.conn.reshape(-1, 8)
        out = np.zeros(c.shape)
        ind = c >= 0
        out[ind] = x[c[ind]]
        for e, a in zip(*np.nonzero(~ind)):
            d = -c[e, a] - 1
            lo, hi = self.dep_ptr[d], self.dep_ptr[d + 1]
            out[e, a] = float(np.dot(self.dep_w[lo:hi], x[self.dep_conn[lo:hi]]))
        return out

    def gp_density(self, x):
        """rho at the 8 Gauss points: trilinear interpolation of the corner
        values (TMROctConstitutive::evalDesignFieldValue,
        src/topology/TMROctConstitutive.cpp:1470-1499), shape (nelem, 8)."""
        return self.corner_values(x) @ gauss_shape().T

    def gather_corners(self, t):
        """sum over elements of per-corner values t (nelem, 8) onto the
        independent nodes (hanging corners to their parents with the
        hanging weights)."""
        c = self.conn.reshape(-1, 8)
        g = np.zeros(self.nnodes)
        ind = c >= 0
        np.add.at(g, c[ind], t[ind])
        for e, a in zip(*np.nonzero(~ind)):
            d = -c[e, a] - 1
            lo, hi = self.dep_ptr[d], self.dep_ptr[d + 1]
            np.add.at(g, self.dep_conn[lo:hi], self.dep_w[lo:hi] * t[e, a])
        return g

    def energy(self, u):
        u = np.ascontiguousarray(u, dtype=np.double)
        e = np.zeros(self.nelem)
        lib().orc_elem_energy(self.nelem, _p(self.conn), _p(self.dep_ptr), _p(self.dep_conn),
                              _p(self.dep_w), _p(self.hdim), _p(self.bcmask), self.E, self.nu,
                              _p(u), _p(e))
        return e


def gauss_shape():
    """N[q, a]: trilinear shape function of corner a (x fastest) at Gauss
    point q (bit d of q set: coordinate d at +1/sqrt(3))."""
    gp = 1.0 / np.sqrt(3.0)
    N = np.zeros((8, 8))
    for q in range(8):
        pt = [gp if (q >> d) & 1 else -gp for d in range(3)]
        for a in range(8):
            v = 1.0
            for d in range(3):
                s = 1.0 if (a >> d) & 1 else -1.0
                v *= 0.5 * (1.0 + s * pt[d])
            N[q, a] = v
    return N


def weight_transpose(rowp, cols, w, xf, ncoarse):
    """TACSBVecInterp::multWeightTranspose as TMRConformFilter::setDesignVars
    uses it (src/topology/TMRConformFilter.cpp:154-167): with P the
    prolongation (rows fine nodes, columns coarse nodes),
    x_c = (P^T x_f) / (P^T 1)."""
    rowp = np.asarray(rowp)
    rows = np.repeat(np.arange(len(rowp) - 1), np.diff(rowp))
    num = np.zeros(ncoarse)
    den = np.zeros(ncoarse)
    np.add.at(num, cols, w * np.asarray(xf)[rows])
    np.add.at(den, cols, w)
    return num / den


def stress_pnorm_gp(orc, rho_gp, u, P=8.0, ys=1.0, eps=0.3, grad=True):
    """Nodal-design p-norm: the relaxation uses the Gauss-point densities;
    returns (f, dfdu, dfdrho_gp (nelem, 8))."""
    L = _stress_sigs()
    rho_gp = np.ascontiguousarray(rho_gp, dtype=np.double).reshape(-1)
    u = np.ascontiguousarray(u, dtype=np.double)
    dfdu = np.zeros(3 * orc.nnodes)
    dr = np.zeros(8 * orc.nelem)
    f = L.orc_stress_pnorm_gp(orc.nnodes, orc.nelem, _p(orc.conn), _p(orc.dep_ptr),
                              _p(orc.dep_conn), _p(orc.dep_w), _p(orc.hdim), _p(orc.bcmask),
                              orc.E, orc.nu, ys, eps, P, _p(rho_gp), _p(u),
                              _p(dfdu) if grad else None, _p(dr) if grad else None)
    return f, dfdu, dr.reshape(-1, 8)


def bcmask_from_forest(forest, bcs):
    """Clamped-dof bit mask per independent node from (name, mask) pairs."""
    n = int(forest.getNodeRange()[1])
    m = np.zeros(n, np.uint8)
    for name, bits in bcs:
        nodes = forest.getNodesWithName(name)
        nodes = nodes[nodes >= 0]
        m[nodes] |= np.uint8(bits)
    return m


def from_forest(forest, bcs, E=1.0, nu=0.3):
    ptr, dc, dw = forest.getDepNodeConn()
    return Oracle(forest.getMeshConn(), ptr, dc, dw, forest.getElementExtents(),
                  bcmask_from_forest(forest, bcs), E, nu)


# ---------------------------------------------------------------------------
# forest oracle (geometric, box lattices)
OCT_DT = np.dtype([("block", np.int32), ("x", np.int32), ("y", np.int32), ("z", np.int32),
                   ("tag", np.int32), ("level", np.int16), ("info", np.int16)])


def _flib():
    L = lib()
    if not hasattr(L, "_forest_sigs"):
        P, I = C.c_void_p, C.c_int
        L.orc_sort_octants.restype = I
        L.orc_sort_octants.argtypes = [P, I, I]
        L.orc_balance.restype = I
        L.orc_balance.argtypes = [I, I, I, P, I, I]
        L.orc_get_leaves.restype = I
        L.orc_get_leaves.argtypes = [P]
        L.orc_is_refined.restype = I
        L.orc_is_refined.argtypes = [P]
        L.orc_create_nodes.restype = I
        L.orc_node_sizes.argtypes = [P, P, P, P]
        L.orc_get_nodes.argtypes = [P, P, P, P, P]
        L.orc_nodes_on_box_face.restype = I
        L.orc_nodes_on_box_face.argtypes = [I, I, P]
        L.orc_save_mesh.argtypes = [I]
        L.orc_interp.restype = I
        L.orc_get_interp.argtypes = [P, P, P]
        L._forest_sigs = True
    return L


def sort_octants(octs, use_node=False):
    a = np.ascontiguousarray(octs, dtype=OCT_DT).copy()
    n = _flib().orc_sort_octants(_p(a), len(a), int(use_node))
    return a[:n]


class OracleMesh:
    """Balanced leaves + nodes of a box-lattice forest, by brute force."""

    def __init__(self, dims, octs, balance_corner=1):
        L = _flib()
        a = np.ascontiguousarray(octs, dtype=OCT_DT)
        n = L.orc_balance(dims[0], dims[1], dims[2], _p(a), len(a), int(balance_corner))
        self.leaves = np.zeros(n, dtype=OCT_DT)
        L.orc_get_leaves(_p(self.leaves))
        self.dims = dims

    def is_refined(self, o):
        a = np.ascontiguousarray(np.array([o], dtype=OCT_DT))
        return bool(_flib().orc_is_refined(_p(a)))

    def create_nodes(self):
        L = _flib()
        self.nindep = L.orc_create_nodes()
        nn, ni, nd, ndc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.orc_node_sizes(C.byref(nn), C.byref(ni), C.byref(nd), C.byref(ndc))
        self.nnodes, self.ndep = nn.value, nd.value
        self.conn = np.zeros((len(self.leaves), 8), np.intc)
        self.dep_ptr = np.zeros(self.ndep + 1, np.intc)
        self.dep_conn = np.zeros(max(ndc.value, 1), np.intc)
        self.dep_w = np.zeros(max(ndc.value, 1))
        self.gxyz = np.zeros((self.nnodes, 3), np.int64)
        L.orc_get_nodes(_p(self.conn), _p(self.dep_ptr), _p(self.dep_conn), _p(self.dep_w),
                        _p(self.gxyz))
        self.dep_conn = self.dep_conn[:ndc.value]
        self.dep_w = self.dep_w[:ndc.value]
        return self

    def nodes_on_box_face(self, axis, side):
        out = np.zeros(self.nnodes, np.intc)
        n = _flib().orc_nodes_on_box_face(axis, side, _p(out))
        return out[:n]

    def save(self, slot):
        _flib().orc_save_mesh(slot)


def interpolation(fine_dims, fine_octs, coarse_octs, balance_corner=1):
    """(rowp, cols, w) of the prolongation coarse -> fine."""
    fm = OracleMesh(fine_dims, fine_octs, balance_corner).create_nodes()
    fm.save(0)
    cm = OracleMesh(fine_dims, coarse_octs, balance_corner).create_nodes()
    cm.save(1)
    L = _flib()
    nnz = L.orc_interp()
    rowp = np.zeros(fm.nindep + 1, np.intc)
    cols = np.zeros(max(nnz, 1), np.intc)
    w = np.zeros(max(nnz, 1))
    L.orc_get_interp(_p(rowp), _p(cols), _p(w))
    return fm, cm, (rowp, cols[:nnz], w[:nnz])


def ref_octant_lib():
    """The reference's own octant sort built in oracle/_ref (None if absent)."""
    path = os.path.join(HERE, "_ref", "libtmr_octant_ref.so")
    if not os.path.exists(path):
        return None
    R = C.CDLL(path)
    R.ref_sort_octants.restype = C.c_int
    R.ref_sort_octants.argtypes = [C.c_void_p, C.c_int, C.c_int]
    R.ref_child_id.restype = C.c_int
    R.ref_child_id.argtypes = [C.c_void_p]
    R.ref_compare.restype = C.c_int
    R.ref_compare.argtypes = [C.c_void_p, C.c_void_p]
    R.ref_sizeof_octant.restype = C.c_int
    return R


def _stress_sigs():
    L = lib()
    if not hasattr(L, "_stress_sigs"):
        D, I, P = C.c_double, C.c_int, C.c_void_p
        L.orc_stress_pnorm.restype = D
        L.orc_stress_pnorm.argtypes = [I, I, P, P, P, P, P, P, D, D, D, D, D, P, P, P, P]
        L._stress_sigs = True
    return L


def stress_pnorm(orc, rho, u, P=8.0, ys=1.0, eps=0.3, grad=False):
    """Value (and, with grad=True, df/du and the explicit df/drho) of the
    p-norm relaxed von Mises functional on an Oracle mesh."""
    L = _stress_sigs()
    rho = np.ascontiguousarray(rho, dtype=np.double)
    u = np.ascontiguousarray(u, dtype=np.double)
    dfdu = np.zeros(3 * orc.nnodes) if grad else None
    dfdr = np.zeros(orc.nelem) if grad else None
    f = L.orc_stress_pnorm(orc.nnodes, orc.nelem, _p(orc.conn), _p(orc.dep_ptr), _p(orc.dep_conn),
                           _p(orc.dep_w), _p(orc.hdim), _p(orc.bcmask), orc.E, orc.nu, ys, eps, P,
                           _p(rho), _p(u), _p(dfdu) if grad else None, _p(dfdr) if grad else None)
    return (f, dfdu, dfdr) if grad else f


def stress_gradient(orc, rho, b, q=3.0, k0=1e-6, P=8.0, ys=1.0, eps=0.3, rtol=1e-13):
    """Full adjoint gradient d f / d rho_e of the p-norm functional (the
    oracle's own CG solves for u and psi)."""
    orc.assemble(np.array([penalty(r, q, k0) for r in rho]))
    u = orc.solve(b, rtol=rtol)
    f, dfdu, dfdr = stress_pnorm(orc, rho, u, P, ys, eps, grad=True)
    psi = orc.solve(dfdu, rtol=rtol)
    prod = orc.energy2(u, psi)
    return f, dfdr - np.array([dpenalty(r, q) for r in rho]) * prod, u, psi


# ---------------------------------------------------------------------------
# nodal design variables (TMROctConformTACSTopoCreator + TMROctConstitutive)
def _pen(r, q, k0, ptype):
    r = np.asarray(r, dtype=np.double)
    return r ** q + k0 if ptype == 0 else r / (1.0 + q * (1.0 - r)) + k0


def _dpen(r, q, ptype):
    r = np.asarray(r, dtype=np.double)
    if ptype == 0:
        return q * r ** (q - 1.0) if q > 1.0 else np.ones_like(r)
    t = 1.0 + q * (1.0 - r)
    return (q + 1.0) / (t * t)


def nodal_assemble(orc, x, q=3.0, k0=1e-6, ptype=0):
    """Assemble K(x) of a nodal design: per Gauss point penalty(rho_q) K_q
    (TMROctConstitutive::evalTangentStiffness, src/topology/TMROctConstitutive.cpp:498-551)."""
    rg = orc.gp_density(x)
    orc.assemble_gp(_pen(rg, q, k0, ptype))
    return rg


def nodal_compliance_gradient(orc, x, b, q=3.0, k0=1e-6, ptype=0, rtol=1e-13):
    """c = b^T u and d c / d x_j = -sum_e sum_q N_a(q) p'(rho_q) u_e^T K_q u_e
    over the corners a of the elements at node j (hanging corners to their
    parents): TMROctConstitutive::addStressDVSens with psi = u
    (src/topology/TMROctConstitutive.cpp:557-654)."""
    rg = nodal_assemble(orc, x, q, k0, ptype)
    u = orc.solve(b, rtol=rtol)
    e = orc.energy_gp(u, u)
    t = (-_dpen(rg, q, ptype) * e) @ gauss_shape()
    return float(b @ u), orc.gather_corners(t), u


def nodal_stress_gradient(orc, x, b, q=3.0, k0=1e-6, ptype=0, P=8.0, ys=1.0, eps=0.3,
                          rtol=1e-13):
    """p-norm value and its gradient per node for a nodal design (relaxation
    at the Gauss-point densities; explicit term plus -psi^T dK/dx u)."""
    rg = nodal_assemble(orc, x, q, k0, ptype)
    u = orc.solve(b, rtol=rtol)
    f, dfdu, dfdr = stress_pnorm_gp(orc, rg, u, P, ys, eps, grad=True)
    psi = orc.solve(dfdu, rtol=rtol)
    e = orc.energy_gp(u, psi)
    t = (dfdr - _dpen(rg, q, ptype) * e) @ gauss_shape()
    return f, orc.gather_corners(t), u, psi


# ---------------------------------------------------------------------------
# multigrid (oracle/mg_oracle.c): Galerkin hierarchy, Chebyshev or symmetric
# SOR smoothing, dense Cholesky at the coarsest level; flexible PCG or
# GMRES(m) around it
def _mglib():
    L = lib()
    if not hasattr(L, "_mg_sigs"):
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.orc_mg_set_level.argtypes = [I, I, P]
        L.orc_mg_set_matrix.argtypes = [I, P, P, P]
        L.orc_mg_set_interp.argtypes = [I, I, I, P, P, P]
        L.orc_mg_galerkin.restype = I
        L.orc_mg_galerkin.argtypes = [I]
        L.orc_mg_get_matrix.argtypes = [I, P, P, P, P, P]
        L.orc_mg_setup.restype = I
        L.orc_mg_setup.argtypes = [I, I, D, D, D]
        L.orc_mg_pcg.restype = I
        L.orc_mg_pcg.argtypes = [P, P, D, I, P]
        L.orc_mg_gmres.restype = I
        L.orc_mg_gmres.argtypes = [P, P, I, D, I, P]
        L.orc_mg_get_lambda.restype = D
        L.orc_mg_get_lambda.argtypes = [I]
        L.orc_mg_set_lambda.restype = I
        L.orc_mg_set_lambda.argtypes = [I, D]
        L.orc_mg_op.restype = I
        L.orc_mg_op.argtypes = [I, I, P, P]
        L.orc_mg_set_coarse.restype = None
        L.orc_mg_set_coarse.argtypes = [I]
        L.orc_mg_set_order.restype = I
        L.orc_mg_set_order.argtypes = [I, I, P]
        L._mg_sigs = True
    return L


class Multigrid:
    """Galerkin multigrid on the oracle's assembled fine operator: levels
    from a finest-first forest list (their interpolation CSRs and clamped
    dofs), A_{l+1} = P~^T A_l P~.  smoother 'chebyshev' (degree, interval as
    fractions of lambda_max(D^-1 A)) or 'sor' (symmetric, omega); the
    coarsest level solved directly or (coarse_direct=False) smoothed once."""

    def __init__(self, orc, forests, bcs, interps=None):
        L = _mglib()
        self.orc, self.nl = orc, len(forests)
        rp, cl, v = orc.csr()
        self.masks = [bcmask_from_forest(f, bcs) for f in forests]
        for l, f in enumerate(forests):
            L.orc_mg_set_level(l, int(f.getNodeRange()[1]), _p(self.masks[l]))
        L.orc_mg_set_matrix(0, _p(rp), _p(cl), _p(v))
        for l in range(self.nl - 1):
            if interps is not None:
                prp, pcl, pw = interps[l]
            else:
                prp, pcl, pw = forests[l].createInterpolation(forests[l + 1])
            prp = np.ascontiguousarray(prp, dtype=np.intc)
            pcl = np.ascontiguousarray(pcl, dtype=np.intc)
            pw = np.ascontiguousarray(pw, dtype=np.double)
            L.orc_mg_set_interp(l, len(prp) - 1, int(forests[l + 1].getNodeRange()[1]),
                                _p(prp), _p(pcl), _p(pw))
            L.orc_mg_galerkin(l)

    def setup(self, smoother="chebyshev", degree=2, lower=1.0 / 6.0, upper=1.1, omega=1.0,
              coarse_direct=True):
        _mglib().orc_mg_set_coarse(1 if coarse_direct else 0)
        r = _mglib().orc_mg_setup(0 if smoother == "chebyshev" else 1, int(degree), lower, upper,
                                  omega)
        if r != 0:
            raise RuntimeError("oracle multigrid setup failed (%d)" % r)

    def set_order(self, l, nodes):
        """SOR sweep order of level l: node indices (each node's 3 dofs in
        turn); None restores row order."""
        if nodes is None:
            r = _mglib().orc_mg_set_order(int(l), 0, None)
        else:
            nodes = np.ascontiguousarray(nodes, dtype=np.intc)
            r = _mglib().orc_mg_set_order(int(l), len(nodes), _p(nodes))
        if r != 0:
            raise ValueError("oracle: bad SOR order for level %d" % l)

    def matrix(self, l):
        import scipy.sparse as sp
        L = _mglib()
        n, nnz = C.c_int(), C.c_int()
        L.orc_mg_get_matrix(l, C.byref(n), C.byref(nnz), None, None, None)
        rp = np.zeros(n.value + 1, np.intc)
        cl = np.zeros(nnz.value, np.intc)
        v = np.zeros(nnz.value)
        L.orc_mg_get_matrix(l, None, None, _p(rp), _p(cl), _p(v))
        return sp.csr_matrix((v, cl, rp), shape=(n.value, n.value))

    def get_lambda(self, l):
        return float(_mglib().orc_mg_get_lambda(int(l)))

    def set_lambda(self, l, lam):
        _mglib().orc_mg_set_lambda(int(l), float(lam))

    OPS = {"apply": 0, "restrict": 1, "prolong": 2, "smooth": 3, "coarse": 4, "vcycle": 5}

    def op(self, name, l, x):
        """One component alone (orc_mg_op), after setup()."""
        code = self.OPS[name]
        nout = 3 * len(self.masks[l + 1 if code == 1 else l])
        x = np.ascontiguousarray(x, dtype=np.double)
        y = np.zeros(nout)
        if _mglib().orc_mg_op(code, int(l), _p(x), _p(y)) != 0:
            raise ValueError("oracle mg op %r on level %d" % (name, l))
        return y

    def _rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.double).copy()
        b[~self.orc.free_mask()] = 0.0
        return b

    def pcg(self, b, rtol=1e-8, maxit=500):
        b = self._rhs(b)
        x = np.zeros_like(b)
        rr = C.c_double()
        it = _mglib().orc_mg_pcg(_p(b), _p(x), rtol, maxit, C.byref(rr))
        self.iters, self.relres = abs(it), rr.value
        if it < 0:
            raise RuntimeError("oracle MG-PCG did not converge")
        return x

    def gmres(self, b, m=50, rtol=1e-8, maxit=1000):
        b = self._rhs(b)
        x = np.zeros_like(b)
        rr = C.c_double()
        it = _mglib().orc_mg_gmres(_p(b), _p(x), int(m), rtol, maxit, C.byref(rr))
        self.iters, self.relres = abs(it), rr.value
        if it < 0:
            raise RuntimeError("oracle MG-GMRES did not converge")
        return x


# ---------------------------------------------------------------------------
# design filters (TMRHelmholtzFilter, TMRMatrixFilter) on assembled scalar
# operators; element matrices as tensor products of the 1D linear element
# (mass h/6 [[2, 1], [1, 2]], stiffness 1/h [[1, -1], [-1, 1]]), a derivation
# independent of the product's quadrature
def scalar_mass_laplacian(forest):
    """(M, K) over the independent nodes, hanging nodes through their
    constraints, scipy CSR."""
    import scipy.sparse as sp
    conn = np.asarray(forest.getMeshConn()).reshape(-1, 8)
    ext = forest.getElementExtents()
    dp, dc, dw = forest.getDepNodeConn()
    n = int(forest.getNodeRange()[1])

    def m1(h):
        return h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])

    def k1(h):
        return 1.0 / h * np.array([[1.0, -1.0], [-1.0, 1.0]])

    rows, cols, mv, kv = [], [], [], []
    idx = [(a & 1, (a >> 1) & 1, a >> 2) for a in range(8)]
    for e in range(len(conn)):
        hx, hy, hz = ext[e]
        mx, my, mz, kx, ky, kz = m1(hx), m1(hy), m1(hz), k1(hx), k1(hy), k1(hz)
        Me = np.zeros((8, 8))
        Ke = np.zeros((8, 8))
        for a, (ia, ja, ka) in enumerate(idx):
            for b, (ib, jb, kb) in enumerate(idx):
                Me[a, b] = mx[ia, ib] * my[ja, jb] * mz[ka, kb]
                Ke[a, b] = (kx[ia, ib] * my[ja, jb] * mz[ka, kb] + mx[ia, ib] * ky[ja, jb] * mz[ka, kb]
                            + mx[ia, ib] * my[ja, jb] * kz[ka, kb])
        ex = []
        for a in range(8):
            g = conn[e, a]
            if g >= 0:
                ex.append([(g, 1.0)])
            else:
                d = -g - 1
                ex.append([(dc[j], dw[j]) for j in range(dp[d], dp[d + 1])])
        for a in range(8):
            for (i, wa) in ex[a]:
                for b in range(8):
                    for (j, wb) in ex[b]:
                        rows.append(i)
                        cols.append(j)
                        mv.append(wa * wb * Me[a, b])
                        kv.append(wa * wb * Ke[a, b])
    M = sp.csr_matrix((mv, (rows, cols)), shape=(n, n))
    K = sp.csr_matrix((kv, (rows, cols)), shape=(n, n))
    return M, K


def helmholtz_filter(M, K, r, x, transpose=False):
    """TMRHelmholtzFilter: x_f = (M + r^2 K)^-1 M x; transpose M (M + r^2 K)^-1 g."""
    import scipy.sparse.linalg as spl
    A = (M + r * r * K).tocsc()
    if not transpose:
        return spl.spsolve(A, M @ x)
    return M @ spl.spsolve(A, x)


def matrix_filter(M, r, N, x, transpose=False):
    """TMRMatrixFilter (src/topology/TMRMatrixFilter.cpp:118-238): D = M 1,
    a = 1 / (1 + r^2 / D^(2/3)), b = a r^2 / (D^(2/3) D), N Horner terms,
    normalised by 1 / F(1)."""
    D = np.asarray(M @ np.ones(M.shape[0])).ravel()
    t = D ** (2.0 / 3.0)
    a = 1.0 / (1.0 + r * r / t)
    b = a * r * r / (t * D)

    def raw(v, tinv):
        t1 = a * v
        out = t1.copy()
        for _ in range(N):
            out = b * (M @ out) + t1
        return tinv * out

    def raw_t(v, tinv):
        t1 = tinv * v
        out = t1.copy()
        for _ in range(N):
            out = M @ (b * out) + t1
        return a * out

    ones = np.ones_like(D)
    y = raw(ones, ones)
    tinv = np.where(y != 0.0, 1.0 / np.where(y != 0.0, y, 1.0), 0.0)
    return raw_t(x, tinv) if transpose else raw(x, tinv)
