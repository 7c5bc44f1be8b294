"""tmrx_problem_stress_pnorm_noalloc (the entry point that allocates
nothing: its partial sums live in a fixed device array of the library)
against tmrx_problem_stress_pnorm, through the C-ABI: two calls in a row, the
second on a smaller problem, each equal to the allocating entry point's value
bit for bit (same kernels, same order of additions), and the first problem
again after the second."""
import ctypes as C

import numpy as np
import pytest

from tmr_amd import _lib, solver, workloads

pytestmark = pytest.mark.gpu


def _solved(forests, rho):
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    prob.set_design(np.clip(rho, 0.05, 1.0))
    prob.solve(prob.traction_load("load", [0.0, 0.0, -1.0]),
               opts=solver.default_opts(rtol=1e-10))
    return prob


def _both(prob, P):
    lib = _lib.lib()
    v_ref, v_na = C.c_double(), C.c_double()
    assert lib.tmrx_problem_stress_pnorm(prob.ptr, P, C.byref(v_ref)) == 0
    r = lib.tmrx_problem_stress_pnorm_noalloc(prob.ptr, P, C.byref(v_na))
    assert r == 0, lib.tmrx_last_error().decode()
    return v_ref.value, v_na.value


def test_noalloc_matches_allocating_entry_point():
    big_forests, big_rho, _ = workloads.adaptive(seed=5, base_level=2, max_level=4, nlevels=3)
    small_forests, small_rho, _ = workloads.cube_l4()
    big = _solved(big_forests, big_rho)
    small = _solved(small_forests, small_rho)
    assert small.nelem < big.nelem
    print("no-alloc p-norm: %d elements, then %d" % (big.nelem, small.nelem), flush=True)
    for prob, P in ((big, 8.0), (small, 6.0)):
        ref, na = _both(prob, P)
        print("  P = %g: allocating %.17g, no-alloc %.17g" % (P, ref, na), flush=True)
        assert np.isfinite(na) and na > 0.0
        assert na == ref
    # the big problem again after the small one (stale partials must not leak)
    ref, na = _both(big, 8.0)
    assert na == ref


def test_noalloc_rejects_bad_arguments():
    forests, rho, _ = workloads.cube_l4()
    prob = _solved(forests, rho)
    lib = _lib.lib()
    v = C.c_double()
    assert lib.tmrx_problem_stress_pnorm_noalloc(prob.ptr, 1.0, C.byref(v)) < 0
    assert "exponent" in lib.tmrx_last_error().decode()
    assert lib.tmrx_problem_stress_pnorm_noalloc(prob.ptr, 8.0, None) < 0
