"""Python handle of the native problem (multigrid hierarchy + FP64 PCG on one
GPU).  Arrays may be numpy (host) or torch tensors on the GPU (device
pointers are passed straight through the C-ABI)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

SIMP = 0
RAMP = 1


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _ptr(a, n=None, name="array", out=False):
    """(pointer, on_device) for a contiguous float64 numpy array or torch
    tensor of n entries (n None: any length).  Anything else raises
    ValueError: the C-ABI reads and writes n doubles from the pointer."""
    if _is_torch(a):
        if str(a.dtype) != "torch.float64":
            raise ValueError("%s: expected a float64 tensor, got %s" % (name, a.dtype))
        if not a.is_contiguous():
            raise ValueError("%s: tensor must be contiguous" % name)
        if n is not None and a.numel() != n:
            raise ValueError("%s: expected %d entries, got %d" % (name, n, a.numel()))
        return C.c_void_p(a.data_ptr()), 1 if a.is_cuda else 0
    if not isinstance(a, np.ndarray):
        raise ValueError("%s: expected a numpy array or a torch tensor" % name)
    if a.dtype != np.float64:
        raise ValueError("%s: expected float64, got %s" % (name, a.dtype))
    if not a.flags.c_contiguous:
        raise ValueError("%s: array must be C-contiguous" % name)
    if out and not a.flags.writeable:
        raise ValueError("%s: output array is read-only" % name)
    if n is not None and a.size != n:
        raise ValueError("%s: expected %d entries, got %d" % (name, n, a.size))
    return C.c_void_p(a.ctypes.data), 0


def _input(a):
    """Host inputs may be any float sequence (converted to contiguous
    float64); tensors are passed as they are and checked by _ptr."""
    if _is_torch(a):
        return a
    return np.ascontiguousarray(a, dtype=np.double)


def _same_side(dev_a, dev_b, what):
    if dev_a != dev_b:
        raise ValueError("%s must both be host arrays or both device tensors" % what)


def default_opts(**kw):
    o = _lib.SolveOpts()
    lib().tmrx_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Problem:
    """Finite-element hierarchy on forests[0] (finest) ... forests[-1]."""

    def __init__(self, forests, E=1.0, nu=0.3, ys=1.0, q=3.0, k0=1e-6, eps=0.3,
                 penalty_type=SIMP, bcs=(), device=0):
        self.forests = list(forests)
        m = _lib.Material(E, nu, ys, q, k0, eps, penalty_type)
        names = [b[0].encode() for b in bcs]
        masks = [int(b[1]) if len(b) > 1 else 7 for b in bcs]
        carr = (C.c_char_p * max(1, len(names)))(*names)
        marr = (C.c_int * max(1, len(masks)))(*masks)
        farr = (C.c_void_p * len(self.forests))(*[f.ptr.value for f in self.forests])
        h = C.c_void_p()
        check(lib().tmrx_problem_create(len(self.forests), farr, C.byref(m), len(names), carr, marr,
                                        int(device), C.byref(h)), "Problem")
        self.ptr = h
        nd, ne, nl = C.c_int(), C.c_int(), C.c_int()
        check(lib().tmrx_problem_get_sizes(self.ptr, C.byref(nd), C.byref(ne), C.byref(nl)))
        self.ndof, self.nelem, self.nlevels = nd.value, ne.value, nl.value
        self.material = dict(E=E, nu=nu, ys=ys, q=q, k0=k0, eps=eps, penalty_type=penalty_type)
        self.info = _lib.SolveInfo()
        self.device = int(device)
        self.generation = 0

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().tmrx_problem_free(self.ptr)
                self.ptr = None
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        check(lib().tmrx_problem_set_stream(self.ptr, C.c_void_p(stream_ptr)))

    def _bump(self):
        """Advance the state generation: every call that changes the design,
        the solution or the installed load case does this, so that a saved
        generation tells whether the state is still the one it saw."""
        self.generation += 1

    def _check_device(self, a):
        if _is_torch(a) and a.is_cuda and a.device.index not in (None, self.device):
            raise ValueError("tensor on cuda:%d, problem on device %d" % (a.device.index, self.device))

    def set_design(self, rho):
        rho = _input(rho)
        self._check_device(rho)
        p, dev = _ptr(rho, self.nelem, "rho")
        check(lib().tmrx_problem_set_design(self.ptr, p, int(self.nelem), dev), "set_design")
        self._bump()

    def set_design_nodal(self, x):
        """One design value per independent node of the finest forest
        (TMROctConformTACSTopoCreator + TMRConformFilter); afterwards the
        gradients are per node."""
        x = _input(x)
        self._check_device(x)
        p, dev = _ptr(x, None, "x")
        n = x.numel() if _is_torch(x) else x.size
        check(lib().tmrx_problem_set_design_nodal(self.ptr, p, int(n), dev), "set_design_nodal")
        self._bump()

    def set_bc_values(self, ubc):
        """Prescribed values of the clamped dofs (one entry per dof; entries of
        free dofs are ignored)."""
        ubc = _input(ubc)
        self._check_device(ubc)
        p, dev = _ptr(ubc, self.ndof, "ubc")
        check(lib().tmrx_problem_set_bc_values(self.ptr, p, int(self.ndof), dev), "set_bc_values")
        self._bump()

    def design_info(self):
        """(mode, gradient length): mode 0 element densities, 1 nodal."""
        m, n = C.c_int(), C.c_int()
        check(lib().tmrx_problem_design_info(self.ptr, C.byref(m), C.byref(n)), "design_info")
        return m.value, n.value

    def level_design(self, level):
        """Nodal design values on multigrid level `level` (0 = finest)."""
        n = self.forests[level].getNodeRange()[1]
        x = np.zeros(n)
        check(lib().tmrx_problem_level_design(self.ptr, int(level),
                                              x.ctypes.data_as(C.POINTER(C.c_double))),
              "level_design")
        return x

    def setup(self, opts=None):
        opts = opts or default_opts()
        check(lib().tmrx_problem_setup(self.ptr, C.byref(opts)), "setup")

    def solve(self, f, u=None, opts=None, allow_unconverged=False):
        """u = K(rho)^-1 f.  With opts.zero_guess = 0, u holds the initial
        guess on entry."""
        opts = opts or default_opts()
        f = _input(f)
        if u is None:
            if not opts.zero_guess:
                raise ValueError("solve: zero_guess = 0 needs the initial guess in u")
            u = f.new_empty(f.shape) if _is_torch(f) else np.zeros(self.ndof)
        self._check_device(f)
        self._check_device(u)
        fp, dev = _ptr(f, self.ndof, "f")
        up, udev = _ptr(u, self.ndof, "u", out=True)
        _same_side(dev, udev, "f and u")
        r = lib().tmrx_problem_solve(self.ptr, fp, up, dev, C.byref(opts), C.byref(self.info))
        self._bump()
        check(r, "solve")
        if r == 1 and not allow_unconverged:
            raise _lib.TmrxError("solve: PCG did not converge (%d iterations, |r|/|b| = %.3e)"
                                 % (self.info.iters, self.info.rnorm / max(self.info.rnorm0, 1e-300)))
        return u

    def solve_multi(self, F, U=None, opts=None, allow_unconverged=False):
        """Solve for every load case (rows of F, shape [nrhs, ndof]) in one
        batched PCG; returns U of the same shape; per-case infos in
        self.infos.  Load case 0 becomes the current state."""
        opts = opts or default_opts()
        F = _input(F)
        if len(F.shape) != 2:
            raise ValueError("F must be [nrhs, ndof]")
        if U is None:
            if not opts.zero_guess:
                raise ValueError("solve_multi: zero_guess = 0 needs the initial guesses in U")
            U = F.new_empty(F.shape) if _is_torch(F) else np.zeros_like(F)
        nrhs = F.shape[0]
        if F.shape[1] != self.ndof:
            raise ValueError("F has %d columns, expected %d" % (F.shape[1], self.ndof))
        self._check_device(F)
        self._check_device(U)
        infos = (_lib.SolveInfo * nrhs)()
        fp, dev = _ptr(F, nrhs * self.ndof, "F")
        up, udev = _ptr(U, nrhs * self.ndof, "U", out=True)
        _same_side(dev, udev, "F and U")
        r = lib().tmrx_problem_solve_multi(self.ptr, nrhs, fp, up, dev, C.byref(opts), infos)
        self._bump()
        check(r, "solve_multi")
        self.infos = list(infos)
        self.info = self.infos[0]
        if r == 1 and not allow_unconverged:
            bad = [k for k, i in enumerate(self.infos) if not i.converged]
            raise _lib.TmrxError("solve_multi: load cases %s did not converge" % bad)
        return U

    def set_multigrid(self, omega=1.0, use_galerkin=False, use_coarse_direct_solve=True,
                      use_chebyshev_smoother=True):
        """TMR_CreateTACSMg's options on this hierarchy (tmrx_mg_create):
        Galerkin coarse operators P^T A P or rediscretised coarse levels; the
        coarsest level solved directly or smoothed once; Chebyshev
        smoothing, or symmetric SOR with relaxation omega
        (use_chebyshev_smoother=False)."""
        check(lib().tmrx_mg_create(self.ptr, float(omega), int(bool(use_galerkin)),
                                   int(bool(use_coarse_direct_solve)),
                                   int(bool(use_chebyshev_smoother))), "set_multigrid")
        self.galerkin = bool(use_galerkin)
        self._bump()

    def smoother_order(self, level):
        """(nodes, color_ptr) of level's SOR sweep after a setup with SOR
        smoothing: the level's nodes in colour-major order, colour c being
        nodes[color_ptr[c]:color_ptr[c + 1]]."""
        n, nc = C.c_int(), C.c_int()
        check(lib().tmrx_problem_smoother_order(self.ptr, int(level), C.byref(n), C.byref(nc),
                                                None, None), "smoother_order")
        nodes = np.zeros(n.value, np.intc)
        cptr = np.zeros(nc.value + 1, np.intc)
        check(lib().tmrx_problem_smoother_order(self.ptr, int(level), None, None,
                                                nodes.ctypes.data_as(_lib.IP),
                                                cptr.ctypes.data_as(_lib.IP)), "smoother_order")
        return nodes, cptr

    def level_matrix(self, level):
        """The Galerkin operator of a coarse level as a scipy BSR matrix
        (3 x 3 blocks; needs set_multigrid(use_galerkin=True) and a setup)."""
        import scipy.sparse as sp
        n, nb = C.c_int(), C.c_int()
        check(lib().tmrx_problem_level_matrix(self.ptr, int(level), C.byref(n), C.byref(nb),
                                              None, None, None), "level_matrix")
        ptr = np.zeros(n.value + 1, np.intc)
        col = np.zeros(nb.value, np.intc)
        val = np.zeros((nb.value, 3, 3))
        check(lib().tmrx_problem_level_matrix(self.ptr, int(level), None, None,
                                              C.c_void_p(ptr.ctypes.data),
                                              C.c_void_p(col.ctypes.data),
                                              C.c_void_p(val.ctypes.data)), "level_matrix")
        return sp.bsr_matrix((val, col, ptr), shape=(3 * n.value, 3 * n.value))

    def num_components(self):
        n = C.c_int()
        check(lib().tmrx_problem_num_components(self.ptr, C.byref(n)))
        return n.value

    def component_info(self, comp, rhs=0):
        """(iterations, converged, true relative residual) of one
        independent component in the last solve."""
        it, cv, rr = C.c_int(), C.c_int(), C.c_double()
        check(lib().tmrx_problem_component_info(self.ptr, int(rhs), int(comp), C.byref(it),
                                                C.byref(cv), C.byref(rr)), "component_info")
        return it.value, bool(cv.value), rr.value

    def tile_candidates(self):
        """Owner-tile sizes worth timing: the powers of two and their
        midpoints from 256 to 1024 nodes, and the sizes that make the tile
        count a whole number of rounds of resident tiles (256 CUs holding 4
        tiles of 512 threads or 8 of 256 threads), which removes the partly
        filled last round on meshes whose node count is not a power of two."""
        n = int(self.forests[0].getNodeRange()[1])
        c = {256, 384, 512, 768, 1024}
        for per_round in (1024, 2048):
            for r in range(1, 9):
                t = -(-n // (r * per_round))
                if 192 <= t <= 1100:
                    c.add(t)
        return tuple(sorted(c))

    def autotune(self, tile_sizes=None, reps=5):
        """Pick the owner-tile size (and FP32 operator form) with the fastest
        finest-level launches of one mixed PCG iteration (three Chebyshev
        steps, a residual, an FP64 residual); returns (chosen size,
        {size: ms}).  Default candidates: tile_candidates()."""
        if tile_sizes is None:
            tile_sizes = self.tile_candidates()
        n = len(tile_sizes)
        ts = (C.c_int * n)(*tile_sizes)
        ms = (C.c_double * n)()
        ch = C.c_int()
        check(lib().tmrx_problem_autotune(self.ptr, ts, n, int(reps), C.byref(ch), ms), "autotune")
        return ch.value, {int(t): (m if m >= 0 else None) for t, m in zip(tile_sizes, ms)}

    FP32_FORMS = ("vector ALUs", "matrix cores (bf16 split)", "matrix cores (FP32)")

    def tuning(self):
        """(finest-level tile size, FP32 operator form): 0 vector ALUs, 1
        matrix cores with bf16 three-way split operands, 2 matrix cores with
        FP32 operands (non-zero: on the matrix cores)."""
        t, m = C.c_int(), C.c_int()
        check(lib().tmrx_problem_tuning(self.ptr, C.byref(t), C.byref(m)))
        return t.value, m.value

    def tile_info(self, level=0):
        """(tile count, nominal tile size) of a level's owner tiles."""
        nt, tn = C.c_int(), C.c_int()
        check(lib().tmrx_problem_tile_info(self.ptr, int(level), C.byref(nt), C.byref(tn)),
              "tile_info")
        return nt.value, tn.value

    def set_tiling(self, tile_nodes):
        """Owner tiles of `tile_nodes` at the finest level (coarse levels by
        the autotuner's rule); overrides the autotuner until its next call."""
        check(lib().tmrx_problem_set_tiling(self.ptr, int(tile_nodes)), "set_tiling")
        self._bump()

    def set_fp32_form(self, form):
        """Pin the FP32 operator form (0, 1, 2 as in tuning()) until the
        next autotune."""
        check(lib().tmrx_problem_set_fp32_form(self.ptr, int(form)), "set_fp32_form")
        self._bump()

    def true_residual(self):
        v = C.c_double()
        check(lib().tmrx_problem_true_residual(self.ptr, C.byref(v)))
        return v.value

    def set_state(self, f, u):
        """Install a stored load case (f, u) as the current state."""
        f, u = _input(f), _input(u)
        self._check_device(f)
        self._check_device(u)
        fp, dev = _ptr(f, self.ndof, "f")
        up, dev2 = _ptr(u, self.ndof, "u")
        _same_side(dev, dev2, "f and u")
        check(lib().tmrx_problem_set_state(self.ptr, fp, up, dev), "set_state")
        self._bump()

    def compliance(self):
        v = C.c_double()
        check(lib().tmrx_problem_compliance(self.ptr, C.byref(v)), "compliance")
        return v.value

    def compliance_grad(self, g=None):
        n = self.design_info()[1]
        if g is None:
            g = np.zeros(n)
        self._check_device(g)
        p, dev = _ptr(g, n, "g", out=True)
        check(lib().tmrx_problem_compliance_grad(self.ptr, p, dev), "compliance_grad")
        return g

    def stress_pnorm(self, P=8.0):
        v = C.c_double()
        check(lib().tmrx_problem_stress_pnorm(self.ptr, float(P), C.byref(v)), "stress_pnorm")
        return v.value

    def stress_pnorm_grad(self, P=8.0, g=None, opts=None):
        opts = opts or default_opts()
        n = self.design_info()[1]
        if g is None:
            g = np.zeros(n)
        self._check_device(g)
        p, dev = _ptr(g, n, "g", out=True)
        info = _lib.SolveInfo()
        r = lib().tmrx_problem_stress_pnorm_grad(self.ptr, float(P), p, dev, C.byref(opts),
                                                 C.byref(info))
        check(r, "stress_pnorm_grad")
        self.adjoint_info = info
        if r == 1:
            raise _lib.TmrxError("stress_pnorm_grad: the adjoint solve did not converge")
        return g

    def apply(self, x, precision=0, y=None):
        x = _input(x)
        if y is None:
            y = np.zeros(self.ndof) if not _is_torch(x) else x.new_empty(x.shape)
        self._check_device(x)
        self._check_device(y)
        xp, dev = _ptr(x, self.ndof, "x")
        yp, ydev = _ptr(y, self.ndof, "y", out=True)
        _same_side(dev, ydev, "x and y")
        check(lib().tmrx_problem_apply(self.ptr, xp, yp, dev, int(precision)), "apply")
        return y

    MG_OPS = {"apply": 0, "restrict": 1, "prolong": 2, "smooth": 3, "coarse": 4, "vcycle": 5}

    def mg_op(self, op, level, x, opts=None):
        """One multigrid component alone on `level` (tmrx_problem_mg_op):
        'apply' A_l x, 'restrict' R_l x (to level + 1), 'prolong' P_l x
        (from level + 1), 'smooth' the pre-smoothing sweep with right-hand
        side x (Chebyshev or symmetric SOR), 'coarse' the coarsest level's
        solve (direct or smoothed), 'vcycle' the V-cycle from
        level.  Runs in the V-cycle's precision (opts.precision).  Returns
        (y, lambda_max of the level)."""
        code = self.MG_OPS[op]
        nl = len(self.forests)
        lin = level + 1 if code == 2 else level
        lout = level + 1 if code == 1 else level
        if not (0 <= lin < nl and 0 <= lout < nl):
            raise ValueError("mg_op: level %d out of range for %r" % (level, op))
        nin = 3 * int(self.forests[lin].getNodeRange()[1])
        nout = 3 * int(self.forests[lout].getNodeRange()[1])
        x = np.ascontiguousarray(x, dtype=np.double)
        if x.shape != (nin,):
            raise ValueError("mg_op: x must have %d entries" % nin)
        y = np.zeros(nout)
        lam = C.c_double()
        opts = opts or default_opts()
        check(lib().tmrx_problem_mg_op(self.ptr, code, int(level), x.ctypes.data, y.ctypes.data, 0,
                                       C.byref(opts), C.byref(lam)), "mg_op")
        return y, lam.value

    def kernel_timing(self, mask=-1):
        """Time the finest-level operator launches whose kind is in `mask`
        (1 product, 2 residual, 4 Chebyshev update, 16 product with dot, -1
        all); False/0 stops."""
        check(lib().tmrx_problem_kernel_timing(self.ptr, int(mask)))

    def kernel_timing_result(self):
        t, n, m = C.c_double(), C.c_int(), C.c_int()
        check(lib().tmrx_problem_kernel_timing_result(self.ptr, C.byref(t), C.byref(n), C.byref(m)))
        return t.value, n.value, m.value

    def traction_load(self, name, tr):
        f = np.zeros(self.ndof)
        t = np.ascontiguousarray(tr, dtype=np.double)
        check(lib().tmrx_problem_traction_load(self.ptr, name.encode(), t.ctypes.data_as(_lib.DP),
                                               f.ctypes.data_as(_lib.DP)), "traction_load")
        return f
