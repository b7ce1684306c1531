"""torch.autograd harness over the solver's adjoint (no reference counterpart).

Objective J(rho) = w_f * f'u + g'u with K(E(rho)) u = f, SIMP modulus
E_e = Emin + rho_e^3 (1 - Emin).  Forward: one MG-PCG solve.  Backward: one
adjoint MG-PCG solve K lambda = w_f f + g and the element kernel
dJ/dE_e = -lambda_e' Ke0 u_e, chained with dE/drho = 3 rho^2 (1 - Emin).
"""
from __future__ import annotations

import numpy as np

EMIN = 1e-9


def simp_modulus(rho: np.ndarray, emin: float = EMIN) -> np.ndarray:
    return emin + rho ** 3 * (1.0 - emin)


def simp_dmodulus(rho: np.ndarray, emin: float = EMIN) -> np.ndarray:
    return 3.0 * rho ** 2 * (1.0 - emin)


def objective_and_gradient(solver, rho: np.ndarray, g: np.ndarray, w_f: float = 1.0):
    """J and dJ/drho (compact element order) through the C ABI."""
    solver.set_modulus(simp_modulus(rho))
    U, info, _ = solver.solve()
    f = solver.query("F")
    J = w_f * float(np.dot(f, U)) + float(np.dot(g, U))
    lam, dJdE, _ = solver.adjoint(w_f * f + g, U)
    return J, dJdE * simp_dmodulus(rho), U


def objective(solver, rho: np.ndarray, g: np.ndarray, w_f: float = 1.0) -> float:
    """J = w_f f'u + g'u alone (no adjoint solve), e.g. for finite differences."""
    solver.set_modulus(simp_modulus(rho))
    U, _, _ = solver.solve()
    return w_f * float(np.dot(solver.query("F"), U)) + float(np.dot(g, U))


def make_function():
    import torch

    class SolverObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, solver, g, w_f):
            r = rho.detach().cpu().double().numpy()
            J, grad, _ = objective_and_gradient(solver, r, g, w_f)
            ctx.save_for_backward(torch.from_numpy(grad).to(rho.device, rho.dtype))
            return torch.tensor(J, dtype=rho.dtype, device=rho.device)

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            return grad_out * grad, None, None, None

    return SolverObjective
