"""Differentiable time evolution: L(theta) = <psi(T)|O|psi(T)> with psi(T) = exp(-i H(theta) T) psi0,
H(theta) = H_0 + sum_p theta_p G_p, evaluated and differentiated on the device through the C-ABI
(te_adjoint_*), and exposed to PyTorch as a torch.autograd.Function.

The forward evolution is the plain Krylov method of the reference (core/krylovTimeEvolver.cpp:
261-512, same step control and error bound); the gradient comes from the adjoint pass of
te_adjoint_run. There is no CPU path: the HIP library must be built.
"""
import numpy as np

from . import capi
from .hamiltonian import as_tuples, createNumberOperator, linInteraction
from .workloads import ring_bose_hubbard_terms


class AdjointProblem:
    """A Hamiltonian family on one basis: terms_fn(theta) gives H(theta)'s term list (in the
    order the reference's builder would see it), generator_terms[p] the term list of G_p =
    dH/dtheta_p, observable_terms the diagonal loss observable O."""

    def __init__(self, ctx, basis, terms_fn, generator_terms, observable_terms, T, tol=1e-8, m=40,
                 checkpoint_interval=0, theta0=None):
        self.ctx = ctx
        self.basis = basis
        self.terms_fn = terms_fn
        self.generators = [capi.operator_build(ctx, basis, as_tuples(g)) for g in generator_terms]
        self.observable = capi.operator_build(ctx, basis, as_tuples(observable_terms))
        self.P = len(self.generators)
        theta0 = np.zeros(self.P) if theta0 is None else np.asarray(theta0, dtype=np.float64)
        self.H = self.hamiltonian(theta0)
        self.adj = capi.Adjoint(ctx, self.H, self.generators, self.observable, T, tol=tol, m=m,
                                checkpoint_interval=checkpoint_interval)

    def hamiltonian(self, theta):
        return capi.operator_build(self.ctx, self.basis, as_tuples(self.terms_fn(np.asarray(theta, dtype=np.float64))))

    def loss_and_grad(self, theta, psi0):
        """(L, dL/dtheta, forward result, adjoint info) at theta."""
        self.H = self.hamiltonian(theta)
        self.adj.set_hamiltonian(self.H)
        return self.adj.run(psi0)


def ring_problem(ctx, L, N, U, T, tol=1e-8, m=40, obs_mode=0, checkpoint_interval=0, theta0=None):
    """Ring Bose-Hubbard with theta = (eps_0..eps_{L-1}, J_0..J_{L-1}):
    H = sum_k eps_k n_k + sum_k J_k (a_k^dag a_{k+1} + h.c.) + (U/2) sum_k n_k (n_k - 1), loss n_obs_mode."""

    def terms_fn(theta):
        return ring_bose_hubbard_terms(L, list(theta[:L]), list(theta[L:]), U)

    gens = [[createNumberOperator(k, 1.0)] for k in range(L)]
    gens += [[linInteraction(k, (k + 1) % L, 0, 0, False, 1.0), linInteraction(k, (k + 1) % L, 0, 0, True, 1.0)]
             for k in range(L)]
    b = capi.Basis.basis(ctx, N, L)
    return AdjointProblem(ctx, b, terms_fn, gens, [createNumberOperator(obs_mode, 1.0)], T, tol=tol, m=m,
                          checkpoint_interval=checkpoint_interval, theta0=theta0)


def _function():
    import torch

    class EvolveLoss(torch.autograd.Function):
        """L(theta) of an AdjointProblem; backward returns grad_output * dL/dtheta."""

        @staticmethod
        def forward(ctx, theta, problem, psi0):
            loss, grad, _, _ = problem.loss_and_grad(theta.detach().cpu().numpy(), psi0)
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype).to(theta.device))
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, grad_output):
            (grad,) = ctx.saved_tensors
            return grad_output * grad, None, None

    return EvolveLoss


_EvolveLoss = None


def evolve_loss(theta, problem, psi0):
    """Differentiable L(theta) (a 0-d tensor of theta's dtype) for a float64 tensor theta."""
    global _EvolveLoss
    if _EvolveLoss is None:
        _EvolveLoss = _function()
    return _EvolveLoss.apply(theta, problem, psi0)
