"""Differentiation through assembly + solve: torch.autograd.Function over libhgdt.

Forward: kappa = exp(log_kappa) per element, assemble A(kappa) and b, solve A u = b (Jacobi-PCG).
Backward: adjoint solve A lambda = dJ/du (A is symmetric; Dirichlet components of dJ/du are
dropped because the constrained rows and columns do not depend on kappa), then the adjoint
element kernel grad_e = -kappa_e lambda_e^T Khat_e u_e, recomputing the unit-coefficient local
matrix element-locally instead of storing local matrices.  No reference counterpart: dune-gdt has
no adjoint path; the result is checked against central finite differences and the CPU oracle.

A DiffusionProblem holds one assembled operator.  Every forward solve gets a new id and keeps its
kappa in the autograd context; a backward whose forward is no longer the one assembled (another solve
on the same problem ran in between) copies its own kappa back and re-assembles before the adjoint
solve, so each graph's gradient is taken at its own kappa.
"""
import numpy as np
import torch

from .gdt import (ContinuousLagrangeSpace, DirichletConstraints, GridFunction, LocalElementIntegralBilinearForm,
                  LocalElementIntegralFunctional, LocalLaplaceIntegrand, LocalProductIntegrand, MatrixOperator,
                  Stencil, VectorFunctional)


class DiffusionProblem:
    """-div(kappa grad u) = f with zero Dirichlet data on a (union) grid, P1/Q1 continuous Lagrange."""

    def __init__(self, grid, source=1.0, rtol=1e-10):
        self.grid = grid
        self.space = ContinuousLagrangeSpace(grid, 1)
        self.kappa = torch.ones(grid.num_elements, dtype=torch.float64, device="cuda")
        self.op = MatrixOperator(grid, self.space, stencil=Stencil.element)
        self.op.append(LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(GridFunction(self.kappa))))
        rhs = VectorFunctional(self.space)
        rhs.append(LocalElementIntegralFunctional(LocalProductIntegrand().with_ansatz(source)))
        self.op.append(rhs)
        self.op.append(DirichletConstraints(self.space))
        self.rtol = rtol
        mask = np.zeros(self.space.num_DoFs, bool)
        mask[self.op.dirichlet_dofs()] = True
        self.free = torch.from_numpy(~mask).to("cuda").to(torch.float64)
        self.last_forward = None
        self.last_adjoint = None
        self._solves = 0      # ids handed to forward solves
        self.assembled = None  # id of the solve whose kappa the operator holds
        self.reassemblies = 0  # backward passes that had to restore their kappa

    def assemble_for(self, kappa, solve_id=None):
        """Assemble A(kappa); returns the id under which the operator now holds it."""
        self.kappa.copy_(kappa)
        self.op.assemble()
        if solve_id is None:
            self._solves += 1
            solve_id = self._solves
        self.assembled = solve_id
        return solve_id

    def solve(self, kappa):
        self.assemble_for(kappa)
        u = self.op.apply_inverse(self.op.rhs, opts={"precision": self.rtol})
        self.last_forward = dict(self.op.last_solve)
        return u


class LogDiffusionToSolution(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_kappa, problem):
        kappa = torch.exp(log_kappa.detach()).to(torch.float64)
        u = problem.solve(kappa)
        ctx.problem = problem
        ctx.solve_id = problem.assembled
        ctx.save_for_backward(u, kappa)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        u, kappa = ctx.saved_tensors
        problem = ctx.problem
        if problem.assembled != ctx.solve_id:
            # another forward solve re-assembled the operator (and overwrote the form's kappa) since this one
            problem.assemble_for(kappa, ctx.solve_id)
            problem.reassemblies += 1
        g = (grad_u.detach().to(torch.float64) * problem.free).contiguous()
        lam = problem.op.apply_inverse(g, opts={"precision": problem.rtol})
        problem.last_adjoint = dict(problem.op.last_solve)
        grad = problem.op.adjoint_log_diffusion_gradient(u, lam)
        return grad, None


def solve_log_diffusion(log_kappa, problem):
    """u(log kappa) with gradients through the adjoint path."""
    return LogDiffusionToSolution.apply(log_kappa, problem)
