"""Gradients of the linear solve through torch.autograd.

torch.autograd.Function is the analogue of feax's jax.custom_vjp around
the linear solve (feax/solvers/linear.py:387-483):

  forward   u = u0 + J_bc^-1 (-R_bc(u0))        (the HIP solve, unchanged)
  backward  lambda = J_bc^-T v                   J_bc^T = J_bc under symmetric
                                                 elimination; lambda[bc] = 0
            grad p    = -(dR/dp)^T lambda        fx_hex8_param_vjp_f64
            grad svar = -(dR/dsvar)^T lambda     fx_hex8_traction_vjp_f64
            grad bc_vals = v_b - (K lambda_f)_b  K the un-eliminated FP64
                                                 Jacobian at u (symmetric for
                                                 the laws here)

The bc_vals cotangent follows feax: -(d R_bc / d bc_vals)^T lambda with
R_bc[b] = u_b - bc_vals gives lambda_b of the full adjoint J^T lambda = v.
Where J keeps the Dirichlet columns (non-symmetric elimination) or the
Newton adjoint corrects lambda_b (feax/solvers/newton.py:200-222,
1031-1069) that is v_b - K_fb^T lambda_f; the linear solve under symmetric
elimination leaves lambda_b = v_b (feax/solvers/linear.py:421-483), its
free DOFs seeing the Dirichlet values only through the initial guess.

The adjoint system reuses the forward's operator: the FP64 CSR (precision
"fp64"; with its multigrid hierarchy under a GeometricMultigrid
preconditioner), or the FP32 matrix inside an FP64 defect correction whose outer
residual is the matrix-free volume residual (precision "mixed"), started
from the Galerkin projection onto the forward's recorded Krylov directions
when they are still held.  Initial guesses get no gradient, as in feax.
"""

import numpy as np


def _bc_leaf(bc):
    """bc.bc_vals when it is a torch tensor that requires grad, else None."""
    import torch
    v = getattr(bc, "bc_vals", None)
    return v if isinstance(v, torch.Tensor) and v.requires_grad else None


def bc_gradient(problem, ts, tp, bc, u, v, lam, coupling):
    """d(v . u)/d bc_vals: v_b, minus (K(u) lambda_f)_b when the Dirichlet
    columns couple (coupling)."""
    import torch
    rows = bc.device(ts.device)[0].long()
    vv = v.detach().to(device=ts.device, dtype=torch.float64).reshape(-1)
    g = vv[rows].clone()
    if coupling:
        from ..assembler import assemble_csr
        lf = lam.detach().to(torch.float64).clone()
        lf[rows] = 0.0
        K = assemble_csr(problem, ts, tp, u.detach(), None, True, torch.float64)
        g -= K.matvec(lf)[rows]
    return g


def _needs_grad(tp):
    import torch
    if tp is None:
        return False
    for v in tuple(tp.volume_vars) + tuple(x for s in tp.surface_vars for x in s):
        if isinstance(v, torch.Tensor) and v.requires_grad:
            return True
    return False


def _leaves(tp):
    """[(path, tensor)] of the torch tensors in a TracedParams; path is
    ("v", i) or ("s", i, j)."""
    import torch
    out = []
    for i, v in enumerate(tp.volume_vars):
        if isinstance(v, torch.Tensor):
            out.append((("v", i), v))
    for i, s in enumerate(tp.surface_vars):
        for j, v in enumerate(s):
            if isinstance(v, torch.Tensor):
                out.append((("s", i, j), v))
    return out


def _rebuild(tp, paths, tensors):
    from ..traced_params import TracedParams
    if tp is None:
        return None
    vv = list(tp.volume_vars)
    sv = [list(s) for s in tp.surface_vars]
    for path, t in zip(paths, tensors):
        if path[0] == "v":
            vv[path[1]] = t
        else:
            sv[path[1]][path[2]] = t
    return TracedParams(volume_vars=tuple(vv), surface_vars=[tuple(s) for s in sv])


def adjoint_solve(problem, ts, tp, bc, v, keep, opts):
    """lambda = J_bc^-1 v with lambda[bc] = 0, on the forward's operator."""
    import torch
    from ..assembler import volume_residual
    from ..csr import transpose_with_maps
    from .common import device_norm, inner_tolerance, krylov, pcg
    rhs = v.detach().to(device=ts.device, dtype=torch.float64).reshape(-1).clone()
    if bc is not None and len(bc.bc_rows):
        rhs.index_fill_(0, bc.device(ts.device)[0].long(), 0.0)
    A, dinv = keep["A"], keep["dinv"]
    from .options import DirectSolverOptions
    if isinstance(opts, DirectSolverOptions) or keep["form"] == "direct":
        from .direct import BandCholesky, direct_adjoint
        if keep["form"] != "direct":
            # a direct adjoint of an iterative forward: factor the FP64 J_bc
            if A is None or A.data.dtype != torch.float64 or A.form != "csr":
                from ..assembler import assemble_csr
                A = assemble_csr(problem, ts, tp, keep.get("u0"), bc, True, torch.float64)
            keep.update(form="direct", A=A, factor=BandCholesky(problem, ts.device).factor(A))
        if isinstance(opts, DirectSolverOptions):
            keep["adjoint_iters"] = [0]
            return direct_adjoint(problem, ts, keep, rhs)
        # an iterative adjoint of a direct forward: Krylov on the kept J_bc
        if opts.use_jacobi_preconditioner:
            from .common import create_jacobi_preconditioner
            dinv = create_jacobi_preconditioner(A, opts.jacobi_shift)
        lam, info = krylov(A, rhs, None, opts, dinv)
        keep["adjoint_iters"] = [info.iters]
        return lam
    if keep["form"] == "matfree":
        from .common import pcg_matfree
        lam, info = pcg_matfree(problem, ts, tp, None, bc, rhs, None, opts.tol, opts.atol, opts.maxiter,
                                opts.check_every)
        keep["adjoint_iters"] = [info.iters]
        return lam
    if keep["form"] == "gmg":
        lam, info = keep["gmg"].solve(rhs, None, opts.tol, opts.atol, opts.maxiter)
        keep["adjoint_iters"] = [info.iters]
        return lam
    if keep["form"] == "fp64":
        if opts.solver in ("bicgstab", "gmres"):
            # J_bc^T: the matrix itself under symmetric elimination, else its
            # transpose through the precomputed maps (feax/solvers/linear.py:
            # 442-456, feax/csr.py:130-141).
            At = A if keep.get("symmetric", True) else transpose_with_maps(
                A, problem.csr_T_perm, problem.csr_T_indptr, problem.csr_T_indices)
            lam, info = krylov(At, rhs, None, opts, dinv)
        else:
            lam, info = pcg(A, rhs, None, opts.tol, opts.atol, opts.maxiter, dinv, None, opts.check_every)
        keep["adjoint_iters"] = [info.iters]
        return lam
    # FP32 operator in an FP64 defect correction.
    rec = keep.get("record")
    if rec is not None and (getattr(rec, "owner", None) is not keep.get("token") or rec.count == 0):
        rec = None
    lam = torch.zeros_like(rhs)
    r = rhs.clone()
    d = torch.empty_like(rhs)
    bnorm = device_norm(r)
    target = max(opts.tol * bnorm, opts.atol)
    rn, outer, budget, inner = bnorm, 0, opts.maxiter, []
    while rn > target and outer < opts.max_outer and budget > 0:
        x0 = rec.project(r) if rec is not None else None
        _, info = pcg(A, r, x0, inner_tolerance(opts, target, rn), 0.0, budget, dinv, True, opts.check_every,
                      out=d, vec_dtype=torch.float64)
        lam += d
        res = volume_residual(problem, ts, tp, lam, bc, negate=True)   # -K_bc lam
        torch.add(rhs, res, out=r)
        rn = device_norm(r)
        inner.append(info.iters)
        budget -= info.iters
        outer += 1
    keep["adjoint_iters"] = inner
    return lam


def parameter_gradients(problem, ts, tp, bc, u, lam, paths, shapes, needs):
    """-(dR_bc/dp)^T lambda for every traced tensor (None where not needed)."""
    from ..assembler import param_vjp, traction_vjp
    grads = []
    tv = None
    for path, shape, need in zip(paths, shapes, needs):
        if not need:
            grads.append(None)
            continue
        if path == ("v", 0):
            g = param_vjp(problem, ts, tp, u, lam, bc, scale=-1.0)
            grads.append(g.reshape(shape))
        elif path[0] == "s" and path[2] == 0:
            if tv is None:
                tv = traction_vjp(problem, ts, lam, bc, scale=-1.0)
            g = tv[path[1]]
            nel = int(np.prod(shape)) if len(shape) else 1
            if nel == g.numel():
                grads.append(g.reshape(shape))
            elif nel == g.shape[0]:
                grads.append(g.sum(dim=1).reshape(shape))
            elif nel == 1:
                grads.append(g.sum().reshape(shape))
            else:
                raise NotImplementedError(f"gradient for a surface var of shape {tuple(shape)}")
        elif path == ("v", 1) and getattr(problem.get_tensor_map(), "alpha", None) is not None:
            # the temperature of a thermo-elastic law
            from ..assembler import thermal_vjp
            grads.append(thermal_vjp(problem, ts, tp, lam, bc, scale=-1.0).reshape(shape))
        else:
            # Further volume/surface vars do not enter the HIP residual.
            import torch
            grads.append(torch.zeros(shape, dtype=torch.float64, device=ts.device))
    return grads


def _function():
    import torch

    class LinearSolveFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, call, *tensors):
            n = len(call["paths"])
            tp = _rebuild(call["tp"], call["paths"], [t.detach() for t in tensors[:n]])
            keep = {}
            u = call["forward"](tp, call["ig"], call["bc"], call["ts"], keep=keep)
            ctx.call, ctx.keep, ctx.tp = call, keep, tp
            ctx.shapes = [tuple(t.shape) for t in tensors]
            ctx.dtypes = [(t.dtype, t.device) for t in tensors]
            ctx.save_for_backward(u)
            return u

        @staticmethod
        def backward(ctx, v):
            (u,) = ctx.saved_tensors
            call, keep = ctx.call, ctx.keep
            ts = keep["ts"]
            n = len(call["paths"])
            needs = ctx.needs_input_grad[1:]
            want_bc = call["has_bc"] and needs[n]
            lam = None
            if any(needs[:n]) or (want_bc and call["bc_coupling"]):
                adj = call["adjoint_fn"] or adjoint_solve
                lam = adj(call["problem"], ts, ctx.tp, call["bc"], v, keep, call["adj_opts"])
            grads = parameter_gradients(call["problem"], ts, ctx.tp, call["bc"], u, lam, call["paths"],
                                        ctx.shapes[:n], needs[:n])
            if call["has_bc"]:
                grads.append(bc_gradient(call["problem"], ts, ctx.tp, call["bc"], u, v, lam, call["bc_coupling"])
                             if want_bc else None)
            grads = [g if g is None else g.to(device=d, dtype=dt) for g, (dt, d) in zip(grads, ctx.dtypes)]
            return (None, *grads)

    return LinearSolveFunction


_FN = None


def differentiable_solve(problem, forward, tp, ig, bc, ts, adj_opts, symmetric, adjoint_fn=None, bc_coupling=True):
    """u = forward(tp, ig, bc, ts), differentiable in the torch tensors of tp
    and in bc.bc_vals when that is a tensor requiring grad.
    adjoint_fn(problem, ts, tp, bc, v, keep, opts) -> lambda (default: the
    linear solve's adjoint_solve on the operator the forward kept);
    bc_coupling: whether the bc_vals gradient carries -(K lambda_f)_b."""
    bc_leaf = _bc_leaf(bc)
    if not _needs_grad(tp) and bc_leaf is None:
        return forward(tp, ig, bc, ts)
    global _FN
    if _FN is None:
        _FN = _function()
    leaves = _leaves(tp) if tp is not None else []
    call = dict(problem=problem, forward=forward, tp=tp, ig=ig, bc=bc, ts=ts, adj_opts=adj_opts,
                paths=[p for p, _ in leaves], adjoint_fn=adjoint_fn, has_bc=bc_leaf is not None,
                bc_coupling=bool(bc_coupling))
    tensors = [t for _, t in leaves] + ([bc_leaf] if bc_leaf is not None else [])
    return _FN.apply(call, *tensors)


def differentiable_linear_solve(problem, forward, tp, ig, bc, ts, adj_opts, symmetric):
    # Under symmetric elimination the forward's free DOFs see the Dirichlet
    # values only through the initial guess, so the bc_vals gradient is v_b.
    return differentiable_solve(problem, forward, tp, ig, bc, ts, adj_opts, symmetric, bc_coupling=not symmetric)


__all__ = ["differentiable_solve", "differentiable_linear_solve", "adjoint_solve", "parameter_gradients",
           "bc_gradient"]
