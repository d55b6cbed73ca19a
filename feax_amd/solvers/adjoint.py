"""Gradients of the linear solve through torch.autograd.

torch.autograd.Function is the analogue of feax's jax.custom_vjp
(feax/solvers/linear.py:387-483).  When no traced parameter needs a
gradient the forward runs alone.
"""


def _needs_grad(tp):
    import torch
    if tp is None:
        return False
    for v in tuple(tp.volume_vars) + tuple(x for s in tp.surface_vars for x in s):
        if isinstance(v, torch.Tensor) and v.requires_grad:
            return True
    return False


def differentiable_linear_solve(problem, forward, tp, ig, bc, ts, adj_opts, symmetric):
    if not _needs_grad(tp):
        return forward(tp, ig, bc, ts)
    raise NotImplementedError("adjoint kernels are not built yet")
