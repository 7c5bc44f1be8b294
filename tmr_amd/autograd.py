"""torch.autograd.Function front end: element densities -> compliance or the
p-norm von Mises functional, with the hand-written adjoint kernels as the
backward (TMRTopoProblem::evalObjConGradient, src/topology/TMRTopoProblem.cpp:1216-1295).

    prob = solver.Problem(forests, ...)
    rho = torch.tensor(..., device="cuda", dtype=torch.float64, requires_grad=True)
    c = compliance(prob, rho, f)        # forward solve on the GPU
    c.backward()                        # rho.grad = d c / d rho (device)

Each forward keeps its own design, load and displacement.  A backward whose
problem has moved on (another forward, solve or set_design in between)
reinstalls them first, so interleaved forwards on one problem each get the
gradient of their own state.
"""
import torch

from . import solver


def _save(ctx, prob, r, f, u):
    ctx.prob = prob
    ctx.save_for_backward(r, f, u)
    ctx.generation = prob.generation


def _restore(ctx):
    """The forward's design and state, reinstalled if anything changed them."""
    prob = ctx.prob
    if prob.generation != ctx.generation:
        r, f, u = ctx.saved_tensors
        prob.set_design(r)
        prob.set_state(f, u)
        ctx.generation = prob.generation
    return prob


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, prob, f, opts):
        r = rho.detach().clone().contiguous()
        prob.set_design(r)
        u = torch.empty_like(f)
        prob.solve(f, u, opts=opts)
        _save(ctx, prob, r, f, u)
        return torch.tensor(prob.compliance(), dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, gout):
        prob = _restore(ctx)
        g = torch.empty(prob.nelem, dtype=torch.float64, device=gout.device)
        prob.compliance_grad(g)
        return gout * g, None, None, None


class _StressPNorm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, prob, f, P, opts):
        r = rho.detach().clone().contiguous()
        prob.set_design(r)
        u = torch.empty_like(f)
        prob.solve(f, u, opts=opts)
        _save(ctx, prob, r, f, u)
        ctx.P, ctx.opts = P, opts
        return torch.tensor(prob.stress_pnorm(P), dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, gout):
        prob = _restore(ctx)
        g = torch.empty(prob.nelem, dtype=torch.float64, device=gout.device)
        prob.stress_pnorm_grad(ctx.P, g, opts=ctx.opts)
        return gout * g, None, None, None, None


def _as_device(f, like):
    if isinstance(f, torch.Tensor):
        return f.to(device=like.device, dtype=torch.float64).contiguous()
    return torch.as_tensor(f, dtype=torch.float64, device=like.device).contiguous()


def compliance(prob, rho, f, opts=None):
    """f^T u(rho) with u = K(rho)^-1 f; differentiable in rho."""
    opts = opts or solver.default_opts(rtol=1e-10)
    return _Compliance.apply(rho, prob, _as_device(f, rho), opts)


def stress_pnorm(prob, rho, f, P=8.0, opts=None):
    """p-norm of the relaxed von Mises failure; differentiable in rho (the
    backward runs the adjoint solve)."""
    opts = opts or solver.default_opts(rtol=1e-10)
    return _StressPNorm.apply(rho, prob, _as_device(f, rho), float(P), opts)
