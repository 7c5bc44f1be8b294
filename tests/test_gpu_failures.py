"""Failure paths and solver options on the GPU: invalid densities are
rejected (host and device arrays), a problem without boundary conditions
fails its coarse factorisation with a clear error, wrong arrays never reach
the C-ABI, and a warm start (zero_guess = 0) converges to the cold-start
solution."""
import numpy as np
import pytest

from tmr_amd import _lib, solver, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube():
    forests, rho, L = workloads.cube_l4()
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[("fixed", 7)])
    prob.set_design(rho)
    return forests, rho, prob


@pytest.mark.parametrize("bad", [np.nan, -0.5, np.inf])
def test_invalid_density_rejected(cube, bad):
    import torch
    forests, rho, prob = cube
    r = rho.copy()
    r[17] = bad
    print("invalid density", bad, flush=True)
    with pytest.raises(_lib.TmrxError, match="design values"):
        prob.set_design(r)
    with pytest.raises(_lib.TmrxError, match="design values"):
        prob.set_design(torch.tensor(r, dtype=torch.float64, device="cuda"))
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    prob.solve(f, opts=solver.default_opts(rtol=1e-10))
    assert prob.true_residual() < 1e-9


def test_wrong_arrays_rejected(cube):
    import torch
    forests, rho, prob = cube
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    with pytest.raises(ValueError, match="float64"):
        prob.solve(torch.tensor(f, dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError, match="entries"):
        prob.solve(f[:-3])
    with pytest.raises(ValueError, match="both"):
        prob.solve(f, torch.zeros(prob.ndof, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="float64"):
        prob.compliance_grad(np.zeros(prob.nelem, dtype=np.float32))


def test_missing_boundary_conditions_fail_setup():
    forests, rho, L = workloads.cube_l4()
    prob = solver.Problem(forests, q=3.0, k0=1e-6, bcs=[])
    prob.set_design(rho)
    with pytest.raises(_lib.TmrxError, match="no clamped dof"):
        prob.setup(solver.default_opts())


@pytest.mark.parametrize("precision", [0, 1])
def test_warm_start(cube, precision):
    """zero_guess = 0: u on entry is the initial guess.  From the solution
    itself no iteration is needed; from a perturbed solution the result
    matches the cold start."""
    forests, rho, prob = cube
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    cold = solver.default_opts(rtol=1e-11, precision=precision)
    u = prob.solve(f, opts=cold)
    it_cold = prob.info.iters
    warm = solver.default_opts(rtol=1e-11, precision=precision, zero_guess=0)
    u2 = u * (1.0 + 1e-3 * np.random.default_rng(1).standard_normal(prob.ndof))
    u3 = prob.solve(f, u2.copy(), opts=warm)
    print("precision %d: cold %d iterations, warm %d" % (precision, it_cold, prob.info.iters),
          flush=True)
    assert np.linalg.norm(u3 - u) / np.linalg.norm(u) < 1e-8
    assert prob.info.iters <= it_cold + 2
    with pytest.raises(ValueError, match="initial guess"):
        prob.solve(f, opts=warm)
