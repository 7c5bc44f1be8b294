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
