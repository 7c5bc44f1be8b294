"""GPU p-norm von Mises functional and its adjoint gradient against the FP64
oracle, central differences, and the torch.autograd.Function wrappers."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import solver, workloads

pytestmark = pytest.mark.gpu

BCS = [("fixed", 7)]
Q, K0 = 3.0, 1e-6


@pytest.fixture(scope="module")
def adapt():
    forests, rho, L = workloads.adaptive(seed=5, base_level=2, max_level=4, nlevels=3)
    rho = np.clip(rho, 0.05, 1.0)
    prob = solver.Problem(forests, q=Q, k0=K0, bcs=BCS, ys=1.0, eps=0.3)
    return forests, rho, prob


@pytest.mark.parametrize("precision", [0, 1])
def test_stress_value_and_gradient_match_oracle(adapt, precision):
    forests, rho, prob = adapt
    print("stress parity precision", precision, "nelem", prob.nelem, flush=True)
    prob.set_design(rho)
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    opts = solver.default_opts(rtol=1e-11, precision=precision)
    prob.solve(f, opts=opts)
    val = prob.stress_pnorm(8.0)
    g = prob.stress_pnorm_grad(8.0, opts=opts)
    orc = po.from_forest(forests[0], BCS)
    val_ref, g_ref, u_ref, psi_ref = po.stress_gradient(orc, rho, f.copy(), q=Q, k0=K0, P=8.0)
    print("value %.12g ref %.12g" % (val, val_ref), flush=True)
    assert abs(val - val_ref) / abs(val_ref) < 1e-6
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("grad rel err %.3e, adjoint true residual %.2e" % (err, prob.adjoint_info.rnorm / prob.adjoint_info.rnorm0), flush=True)
    assert err < 1e-6


def test_stress_gradient_central_differences(adapt):
    forests, rho, prob = adapt
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    opts = solver.default_opts(rtol=1e-12, precision=0)
    prob.set_design(rho)
    prob.solve(f, opts=opts)
    g = prob.stress_pnorm_grad(8.0, opts=opts)
    rng = np.random.default_rng(3)
    h = 1e-6
    for k in rng.choice(prob.nelem, 4, replace=False):
        vals = []
        for sgn in (1, -1):
            r2 = rho.copy()
            r2[k] += sgn * h
            prob.set_design(r2)
            prob.solve(f, opts=opts)
            vals.append(prob.stress_pnorm(8.0))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("elem", k, "adjoint %.8e fd %.8e" % (g[k], fd), flush=True)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(vals[0]))
    prob.set_design(rho)


def test_autograd_functions(adapt):
    import torch
    from tmr_amd import autograd as ag
    forests, rho, prob = adapt
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    r = torch.tensor(rho, dtype=torch.float64, device="cuda", requires_grad=True)
    c = ag.compliance(prob, r, f)
    c.backward()
    g_c = r.grad.detach().cpu().numpy().copy()
    prob.set_design(rho)
    prob.solve(f, opts=solver.default_opts(rtol=1e-10))
    assert float(c) == pytest.approx(prob.compliance(), rel=1e-9)
    assert np.allclose(g_c, prob.compliance_grad(), rtol=1e-8, atol=0)
    r.grad = None
    s = ag.stress_pnorm(prob, r, f, P=8.0)
    s.backward()
    g_s = r.grad.detach().cpu().numpy()
    prob.set_design(rho)
    prob.solve(f, opts=solver.default_opts(rtol=1e-10))
    assert float(s) == pytest.approx(prob.stress_pnorm(8.0), rel=1e-9)
    assert np.allclose(g_s, prob.stress_pnorm_grad(8.0), rtol=1e-6, atol=1e-12)
