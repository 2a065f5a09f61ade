"""Differentiable FETI solve: compliance and its gradient with respect to the
per-element Young's modulus, behind torch.autograd.

forward: assemble the structured elasticity problem with E_e from the tensor,
solve with the MI355X FETI core, return c = f^T u.
backward: adjoint solve K w = dc/du (= f for the compliance) with
espreso_feti_resolve, which reuses the assembled F_i, S_D, interior factors and
coarse space; then dc/dE_e = -w_e^T K_e0 u_e (K_e = E_e K_e0) on the GPU.

PyTorch is only the plumbing here: the tensors are copied to host arrays for
the C ABI.  This extends the reference, which has no adjoint.
"""
import numpy as np
import torch

from . import core


class FETICompliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, young, spec, options):
        E = young.detach().to("cpu", torch.float64).numpy()
        solver = core.GridFETI(dict(spec, material_elem=E), **options)
        try:
            solver.setup()
            solver.solve()
            ctx.u = solver.nodal_solution()
            c = solver.load_work()
        except Exception:
            solver.close()
            raise
        ctx.solver = solver
        return torch.tensor(c, dtype=young.dtype, device=young.device)

    @staticmethod
    def backward(ctx, grad_c):
        solver = ctx.solver
        try:
            # adjoint load dc/du = f; the second solve reuses the factorised operators
            solver.set_nodal_rhs(solver.nodal_load())
            solver.resolve()
            w = solver.nodal_solution()
            sens = solver.element_sensitivity(ctx.u, w)
        finally:
            solver.close()
        grad = torch.from_numpy(sens).to(device=grad_c.device, dtype=grad_c.dtype) * grad_c
        return grad, None, None


def compliance(young, spec, **options):
    """c = f^T u for per-element Young's moduli `young` (tensor, one per element,
    index ex + NEx*(ey + NEy*ez)); differentiable with respect to `young`.

    The adjoint K w = f reuses the forward problem's constraints, which is the
    adjoint of the compliance only when the prescribed values are zero; specs
    with non-zero Dirichlet values are refused."""
    fixed = spec.get("fixed", (0,) * 6)
    values = spec.get("fixed_value", (0.0,) * 6)
    if any(f and v != 0.0 for f, v in zip(fixed, values)):
        raise ValueError("compliance gradient: non-zero prescribed values (fixed_value) are not supported; "
                         "the adjoint solve assumes homogeneous Dirichlet conditions")
    if spec.get("physics", 0) != 0:
        raise ValueError("compliance gradient: elasticity only (per-element Young's modulus)")
    return FETICompliance.apply(young, spec, options)

