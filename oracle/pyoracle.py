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
        _L = L
    return _L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def hex8_stiffness(hx, hy, hz, E=1.0, nu=0.3):
    K = np.zeros((24, 24))
    lib().orc_hex8_stiffness(hx, hy, hz, E, nu, _p(K))
    return K


def penalty(rho, q, k0, ptype=0):
    return lib().orc_penalty(rho, q, k0, ptype)


def dpenalty(rho, q, ptype=0):
    return lib().orc_dpenalty(rho, q, ptype)


class Oracle:
    """Assembled FP64 operator A = Z^T K(rho) Z with clamped dofs."""

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
        return nnz

    def csr(self):
        n = 3 * self.nnodes
        rowp = np.zeros(n + 1, np.intc)
        cols = np.zeros(self.nnz, np.intc)
        vals = np.zeros(self.nnz)
        lib().orc_get_csr(_p(rowp), _p(cols), _p(vals))
        return rowp, cols, vals

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.double)
        y = np.zeros_like(x)
        lib().orc_matvec(_p(x), _p(y))
        return y

    def solve(self, b, rtol=1e-12, maxit=20000):
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

    def energy2(self, u, v):
        u = np.ascontiguousarray(u, dtype=np.double)
        v = np.ascontiguousarray(v, dtype=np.double)
        e = np.zeros(self.nelem)
        lib().orc_elem_energy2(self.nelem, _p(self.conn), _p(self.dep_ptr), _p(self.dep_conn),
                               _p(self.dep_w), _p(self.hdim), _p(self.bcmask), self.E, self.nu,
                               _p(u), _p(v), _p(e))
        return e

    def energy(self, u):
        u = np.ascontiguousarray(u, dtype=np.double)
        e = np.zeros(self.nelem)
        lib().orc_elem_energy(self.nelem, _p(self.conn), _p(self.dep_ptr), _p(self.dep_conn),
                              _p(self.dep_w), _p(self.hdim), _p(self.bcmask), self.E, self.nu,
                              _p(u), _p(e))
        return e


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
