"""Newton solver with Armijo line search (feax/solvers/newton.py:809-1112 API).

create_newton_solver(problem, bc, solver_options, adjoint_solver_options,
newton_options, traced_params, symmetric_elimination, traced_structure)
returns ``solver(traced_params, initial_guess=None, bc=None,
traced_structure=None)``.  One Newton step is

  res = R_bc(sol)                      fx_hex8_residual_f64 (FP64)
  J   = dR_bc/du at sol                fx_hex8_assemble_* (tangent, BC-eliminated)
  du  = PCG(J, -res, x0 = lift(sol))   fx_pcg_*  (x0 as feax/solvers/common.py:23-53)
  sol = Armijo(sol, du)                merit 0.5 ||R_bc||^2, c1, rho, max_backtracks
                                       (feax/solvers/newton.py:809-859)

until ||res|| <= tol, ||res|| / ||res_0|| <= rel_tol, max_iter steps, or a
failed line search (which raises NewtonLineSearchError when
raise_on_line_search_failure, as feax does for eager calls).

precision "fp64": FP64 tangent and PCG.  "mixed": FP32 difference-form
tangent with FP64 PCG vectors inside an FP64 defect correction on the
matrix-free FP64 tangent product, so each Newton direction meets the
linear tolerance of the FP64 tangent; a step whose defect correction stalls
(eps32 * cond(J) near 1, e.g. strongly compressed states) is taken with the
FP64 tangent instead.  The Newton residual and the line search stay FP64.

Gradients (torch.autograd): lambda = J_bc(u*)^-1 v on the FP64 tangent at
the solution, grads = -(dR/dp)^T lambda (fx_hex8_param_vjp_f64); gradients
with respect to Dirichlet values are not provided.
"""

import numpy as np

from .. import _native as nat
from ..assembler import TracedStructure, assemble_csr, residual
from ..DCboundary import DirichletBC
from .common import create_x0, device_norm, inner_tolerance, pcg
from .linear import SolveStats, _dev_vec, _zero_like
from .options import DirectSolverOptions, KrylovSolverOptions, MatrixProperty, NewtonOptions, resolve_iterative_solver


class NewtonLineSearchError(RuntimeError):
    """Armijo backtracking found no descent step (feax/solvers/newton.py:48-57)."""


class NewtonStats(SolveStats):
    def __init__(self):
        super().__init__()
        self.newton_iters = 0
        self.res_history = []
        self.alphas = []
        self.linear_iters = []
        self.fp64_fallbacks = 0

    def __repr__(self):
        return (f"NewtonStats(newton_iters={self.newton_iters}, res={self.res_history}, "
                f"alphas={self.alphas}, linear_iters={self.linear_iters})")


def newton_solve(problem, ts, tp, bc, sol0, lin, newton, symmetric=True, stats=None, keep=None):
    """The Newton loop on the device; returns the FP64 solution vector."""
    import torch
    sol = sol0.clone()
    res = residual(problem, ts, tp, sol, bc)
    r0 = device_norm(res)
    rn = r0
    direct = isinstance(lin, DirectSolverOptions)
    if direct:
        from .direct import BandCholesky
        factor = BandCholesky(problem, ts.device)
        x0_fn, mixed, shift = None, False, None
    else:
        x0_fn = lin.x0_fn if lin.x0_fn is not None else create_x0(bc.bc_rows, bc.bc_vals)
        mixed = lin.precision != "fp64"
        shift = lin.jacobi_shift if lin.use_jacobi_preconditioner else None
        if mixed and problem.vec != 3:
            raise NotImplementedError("mixed precision uses the block-3 difference form (vec = 3)")
    it, found, fallbacks = 0, True, 0
    hist, alphas, lits = [r0], [], []
    if newton.verbose:
        print(f"[feax_amd] Newton start: ||res|| = {r0:.6e}", flush=True)
    trial = torch.empty_like(sol)
    while it < newton.max_iter and rn > newton.tol and rn / max(r0, 1e-30) > newton.rel_tol and found:
        out = assemble_csr(problem, ts, tp, sol, bc, symmetric, torch.float32 if mixed else torch.float64,
                           diff_form=mixed, jacobi_shift=shift)
        J, dinv = out if shift is not None else (out, None)
        b = -res
        du = None
        if direct:
            # feax's direct branch: du = J_bc^-1 b, no initial iterate
            du, n_lin = factor.factor(J).solve(b), 0
        elif mixed:
            du, n_lin = _mixed_direction(problem, ts, tp, bc, sol, b, x0_fn(sol), J, dinv, lin)
            if du is None:
                # The FP32 tangent is too coarse for this state (the defect
                # correction stalls when eps32 * cond(J) approaches 1): take
                # this step with the FP64 tangent.
                fallbacks += 1
                out = assemble_csr(problem, ts, tp, sol, bc, symmetric, torch.float64, jacobi_shift=shift)
                J, dinv = out if shift is not None else (out, None)
                du, info = pcg(J, b, x0_fn(sol), lin.tol, lin.atol, lin.maxiter, dinv, None, lin.check_every)
                n_lin += info.iters
        elif lin.solver in ("bicgstab", "gmres"):
            from .common import krylov
            du, info = krylov(J, b, x0_fn(sol), lin, dinv)
            n_lin = info.iters
        else:
            du, info = pcg(J, b, x0_fn(sol), lin.tol, lin.atol, lin.maxiter, dinv, None, lin.check_every)
            n_lin = info.iters
        lits.append(n_lin)
        # Armijo on 0.5 ||R_bc||^2 with slope -||res||^2.
        grad_merit = -rn * rn
        alpha, found = 1.0, False
        for _ in range(newton.line_search_max_backtracks):
            nat.call("fx_axpby_f64", nat.stream(), sol.numel(), 1.0, nat.ptr(sol), alpha, nat.ptr(du),
                     nat.ptr(trial))
            tres = residual(problem, ts, tp, trial, bc)
            tn = device_norm(tres)
            if np.isfinite(tn) and 0.5 * (tn * tn - rn * rn) <= newton.line_search_c1 * alpha * grad_merit:
                found = True
                sol, trial = trial, sol
                res, rn = tres, tn
                break
            alpha *= newton.line_search_rho
        if not found:
            alpha = 1e-8
            nat.call("fx_axpby_f64", nat.stream(), sol.numel(), 1.0, nat.ptr(sol), alpha, nat.ptr(du),
                     nat.ptr(sol))
            res = residual(problem, ts, tp, sol, bc)
            rn = device_norm(res)
        it += 1
        hist.append(rn)
        alphas.append(alpha)
        if newton.verbose:
            print(f"[feax_amd] Newton iter {it}: ||res|| = {rn:.6e} (rel {rn / max(r0, 1e-30):.3e})", flush=True)
    if stats is not None:
        stats.newton_iters, stats.res_history, stats.alphas, stats.linear_iters = it, hist, alphas, lits
        stats.fp64_fallbacks = fallbacks
        stats.iters = int(sum(lits))
        stats.outer = it
        stats.rs = rn * rn
        stats.relres = rn / r0 if r0 > 0 else 0.0
    if newton.raise_on_line_search_failure and not found:
        raise NewtonLineSearchError(
            f"Newton iter {it}: Armijo line search failed after {newton.line_search_max_backtracks} "
            "backtracks; the Newton direction is not a descent direction for 0.5*||r||^2.")
    if keep is not None:
        keep.update(form="newton", A=None, dinv=None, sol=sol)
    return sol


def _mixed_direction(problem, ts, tp, bc, sol, rhs, x0, A32, dinv, lin):
    """du with J_bc(sol) du = rhs to the linear tolerance: FP32 difference-form
    tangent inside an FP64 defect correction whose residual rhs - J_bc du is
    the matrix-free FP64 tangent product (fx_hex8_jvp_f64); later inner
    solves start from the projection onto the first one's directions."""
    import torch
    from ..assembler import jvp
    from .linear import _record_for
    n = ts.n
    du = x0.clone()
    Kx = jvp(problem, ts, tp, sol, du, bc)
    r = rhs - Kx
    bnorm = device_norm(rhs)
    target = max(lin.tol * bnorm, lin.atol)
    rn = device_norm(r)
    d = torch.empty_like(du)
    record = _record_for(n, min(lin.recycle, lin.maxiter), du.device) if lin.recycle > 0 else None
    if record is not None:
        record.count = 0
        record.owner = object()
    total, outer, budget = 0, 0, lin.maxiter
    while rn > target and outer < lin.max_outer and budget > 0:
        x0i = record.project(r) if (record is not None and outer > 0) else None
        rec = record if (record is not None and outer == 0) else None
        _, info = pcg(A32, r, x0i, inner_tolerance(lin, target, rn), 0.0, budget, dinv, True, lin.check_every,
                      out=d, record=rec)
        du += d
        jvp(problem, ts, tp, sol, du, bc, out=Kx)
        torch.sub(rhs, Kx, out=r)
        prev, rn = rn, device_norm(r)
        total += info.iters
        budget -= info.iters
        outer += 1
        if not rn < 0.5 * prev:
            return None, total   # stalled: the caller takes the FP64 tangent
    return (du, total) if rn <= target else (None, total)


def newton_adjoint(problem, ts, tp, bc, v, keep, opts, symmetric=True):
    """lambda = J_bc(u*)^-1 v (lambda[bc] = 0) on the FP64 tangent at u*."""
    import torch
    sol = keep["sol"]
    direct = isinstance(opts, DirectSolverOptions)
    shift = None if direct or not opts.use_jacobi_preconditioner else opts.jacobi_shift
    out = assemble_csr(problem, ts, tp, sol, bc, symmetric, torch.float64, jacobi_shift=shift)
    J, dinv = out if shift is not None else (out, None)
    rhs = v.detach().to(device=ts.device, dtype=torch.float64).reshape(-1).clone()
    if bc is not None and len(bc.bc_rows):
        rhs.index_fill_(0, bc.device(ts.device)[0].long(), 0.0)
    if direct:
        from .direct import BandCholesky
        keep["adjoint_iters"] = [0]
        return BandCholesky(problem, ts.device).factor(J).solve(rhs)
    if opts.solver in ("bicgstab", "gmres"):
        # J_bc^T: J_bc itself only for the symmetric elimination of a
        # symmetric tangent (not the coupled (u, T) one)
        from .common import krylov
        At = J if (symmetric and not getattr(problem, "coupled", False)) else J.transpose()
        lam, info = krylov(At, rhs, None, opts, dinv)
    else:
        lam, info = pcg(J, rhs, None, opts.tol, opts.atol, opts.maxiter, dinv, None, opts.check_every)
    keep["adjoint_iters"] = [info.iters]
    return lam


def linearized_guess(problem, bc, traced_params, u_ref=None, traced_structure=None, options=None):
    """Initial guess for a displacement-controlled load step: the solution of
    the problem linearised at u_ref (default 0) with the Dirichlet values of
    bc, u = u_ref + w + d with w the BC increment (w_b = bc_vals - u_ref_b)
    and J_bc(u_ref) d = -(R(u_ref) + K(u_ref) w) on the free rows.  Spreads a
    prescribed boundary increment through the body instead of leaving it in
    the boundary element layer (where "previous state with the new BC
    values" puts it)."""
    import torch
    from ..assembler import _ts_or_build
    ts = _ts_or_build(problem, traced_structure)
    opts = options or KrylovSolverOptions(solver="cg", tol=1e-8, atol=0.0, maxiter=20000,
                                          use_jacobi_preconditioner=True)
    u_ref = torch.zeros(ts.n, dtype=torch.float64, device=ts.device) if u_ref is None else _dev_vec(u_ref, ts)
    rows, _, vals = bc.device(ts.device)
    rows = rows.long()
    w = torch.zeros_like(u_ref)
    w[rows] = vals - u_ref[rows]
    K = assemble_csr(problem, ts, traced_params, u_ref, None, True, torch.float64)
    b = -(residual(problem, ts, traced_params, u_ref) + K.matvec(w))
    b[rows] = 0.0
    J, dinv = assemble_csr(problem, ts, traced_params, u_ref, bc, True, torch.float64, jacobi_shift=1e-12)
    d, _ = pcg(J, b, None, opts.tol, opts.atol, opts.maxiter, dinv if opts.use_jacobi_preconditioner else None,
               None, opts.check_every)
    return u_ref + w + d


def create_newton_solver(problem, bc, solver_options=None, adjoint_solver_options=None, newton_options=None,
                         traced_params=None, symmetric_elimination=True, traced_structure=None):
    if solver_options is None:
        solver_options = KrylovSolverOptions(solver="cg", use_jacobi_preconditioner=True)
    for o in (solver_options, adjoint_solver_options):
        if o is not None and not isinstance(o, (KrylovSolverOptions, DirectSolverOptions)):
            raise TypeError(f"Expected DirectSolverOptions or KrylovSolverOptions, got {type(o).__name__}.")
    if isinstance(solver_options, DirectSolverOptions) or isinstance(adjoint_solver_options, DirectSolverOptions):
        if not symmetric_elimination:
            raise NotImplementedError("the direct solver factors the symmetric elimination's SPD J_bc "
                                      "(symmetric_elimination=True)")

    def _resolve(o):
        return o if isinstance(o, DirectSolverOptions) else resolve_iterative_solver(o, MatrixProperty.SPD)
    solver_options = _resolve(solver_options)
    adj = solver_options if adjoint_solver_options is None else _resolve(adjoint_solver_options)
    newton_options = newton_options or NewtonOptions()
    default_bc = bc
    ts_default = traced_structure
    stats = NewtonStats()

    def forward(tp, ig, eff_bc, ts, keep=None):
        if ts is None:
            ts = getattr(problem, "_default_ts", None) or TracedStructure.from_problem(problem)
            problem._default_ts = ts
        if keep is not None:
            keep["ts"] = ts
        u0 = _dev_vec(ig, ts) if ig is not None else _zero_like(problem, eff_bc, ts)
        return newton_solve(problem, ts, tp, eff_bc, u0, solver_options, newton_options, symmetric_elimination,
                            stats, keep)

    def solver(traced_params, initial_guess=None, bc=None, traced_structure=None):
        eff_bc = bc if isinstance(bc, DirichletBC) else default_bc
        ts = traced_structure if traced_structure is not None else ts_default
        from .adjoint import differentiable_solve
        return differentiable_solve(problem, forward, traced_params, initial_guess, eff_bc, ts, adj,
                                    symmetric_elimination,
                                    adjoint_fn=lambda *a: newton_adjoint(*a, symmetric=symmetric_elimination))

    solver.stats = stats
    return solver


__all__ = ["create_newton_solver", "newton_solve", "newton_adjoint", "linearized_guess", "NewtonLineSearchError",
           "NewtonStats"]
