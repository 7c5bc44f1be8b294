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
    """Central differences of the p-norm along a random density direction
    (h = 1e-5: per-element differences at h = 1e-6 drown in the solve
    tolerance, f ~ 5e3), and at the element of largest sensitivity."""
    forests, rho, prob = adapt
    f = prob.traction_load("load", [0.0, 0.0, -1.0])
    opts = solver.default_opts(rtol=1e-12, precision=0)
    prob.set_design(rho)
    prob.solve(f, opts=opts)
    g = prob.stress_pnorm_grad(8.0, opts=opts)

    def fval(r2):
        prob.set_design(r2)
        prob.solve(f, opts=opts)
        return prob.stress_pnorm(8.0)

    d = np.random.default_rng(11).standard_normal(prob.nelem)
    h = 1e-5
    fd = (fval(rho + h * d) - fval(rho - h * d)) / (2 * h)
    print("directional: adjoint %.10e fd %.10e" % (g @ d, fd), flush=True)
    assert abs(fd - g @ d) / abs(g @ d) < 1e-5
    k = int(np.argmax(np.abs(g)))
    e = np.zeros(prob.nelem)
    e[k] = 1.0
    h = 1e-4
    fd = (fval(rho + h * e) - fval(rho - h * e)) / (2 * h)
    print("elem", k, "adjoint %.8e fd %.8e" % (g[k], fd), flush=True)
    assert abs(fd - g[k]) / abs(g[k]) < 1e-3
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
    assert float(c.detach()) == pytest.approx(prob.compliance(), rel=1e-9)
    assert np.allclose(g_c, prob.compliance_grad(), rtol=1e-8, atol=0)
    r.grad = None
    s = ag.stress_pnorm(prob, r, f, P=8.0)
    s.backward()
    g_s = r.grad.detach().cpu().numpy()
    prob.set_design(rho)
    prob.solve(f, opts=solver.default_opts(rtol=1e-10))
    assert float(s.detach()) == pytest.approx(prob.stress_pnorm(8.0), rel=1e-9)
    # same options as the autograd path (rtol 1e-10 for both solves)
    g_ref = prob.stress_pnorm_grad(8.0, opts=solver.default_opts(rtol=1e-10))
    assert np.allclose(g_s, g_ref, rtol=1e-12, atol=0)


def test_autograd_interleaved_forwards(adapt):
    """Two forwards on one problem (different loads, then different designs)
    before their backwards: each gradient is the one of its own state."""
    import torch
    from tmr_amd import autograd as ag
    forests, rho, prob = adapt
    opts = solver.default_opts(rtol=1e-10)
    f1 = prob.traction_load("load", [0.0, 0.0, -1.0])
    f2 = prob.traction_load("load", [0.0, 1.0, 0.0])
    rho2 = np.clip(rho * 0.8 + 0.1, 0.05, 1.0)

    def separate(r_np, f):
        prob.set_design(r_np)
        prob.solve(f, opts=opts)
        return prob.compliance(), prob.compliance_grad().copy()

    c1, g1 = separate(rho, f1)
    c2, g2 = separate(rho, f2)
    c3, g3 = separate(rho2, f1)
    r = torch.tensor(rho, dtype=torch.float64, device="cuda", requires_grad=True)
    (ag.compliance(prob, r, f1, opts) + ag.compliance(prob, r, f2, opts)).backward()
    g = r.grad.detach().cpu().numpy()
    print("interleaved loads: rel err %.2e" % (np.linalg.norm(g - g1 - g2) / np.linalg.norm(g1 + g2)),
          flush=True)
    assert np.allclose(g, g1 + g2, rtol=1e-8, atol=1e-14 * np.abs(g1 + g2).max())
    ra = torch.tensor(rho, dtype=torch.float64, device="cuda", requires_grad=True)
    rb = torch.tensor(rho2, dtype=torch.float64, device="cuda", requires_grad=True)
    ca = ag.compliance(prob, ra, f1, opts)
    cb = ag.compliance(prob, rb, f1, opts)
    assert float(ca.detach()) == pytest.approx(c1, rel=1e-9)
    assert float(cb.detach()) == pytest.approx(c3, rel=1e-9)
    ca.backward()
    cb.backward()
    assert np.allclose(ra.grad.cpu().numpy(), g1, rtol=1e-8, atol=1e-14 * np.abs(g1).max())
    assert np.allclose(rb.grad.cpu().numpy(), g3, rtol=1e-8, atol=1e-14 * np.abs(g3).max())
    ss = ag.stress_pnorm(prob, ra, f2, P=8.0, opts=opts)
    sc = ag.compliance(prob, rb, f1, opts)
    ra.grad = None
    ss.backward()
    prob.set_design(rho)
    prob.solve(f2, opts=opts)
    g_ref = prob.stress_pnorm_grad(8.0, opts=opts)
    assert np.allclose(ra.grad.cpu().numpy(), g_ref, rtol=1e-7, atol=1e-12 * np.abs(g_ref).max())
    rb.grad = None
    sc.backward()
    assert np.allclose(rb.grad.cpu().numpy(), g3, rtol=1e-8, atol=1e-14 * np.abs(g3).max())
    prob.set_design(rho)
