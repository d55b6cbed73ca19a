"""Differentiable linear solve (feax/solvers/linear.py:166-499 API).

create_linear_solver(problem, bc, solver_options, ...) returns
``solver(traced_params, initial_guess=None, bc=None, traced_structure=None)``
computing u = u0 + du with J_bc du = -R_bc(u0) (linear.py:387-417):

* precision "fp64": FP64 CSR assembly with the Dirichlet elimination fused
  into the store, FP64 right-hand side from the residual kernel, HIP
  Jacobi-PCG.
* preconditioner GeometricMultigrid: FP64 CG preconditioned by a V-cycle
  on the structured grid hierarchy (FP32 V-cycle with precision "mixed").
* solver "bicgstab" / "gmres": FP64 CSR, HIP BiCGSTAB or restarted GMRES
  (jax bicgstab / gmres semantics): the methods for the non-symmetric
  operator of symmetric_elimination=False.
* precision "mixed": FP32 CSR assembly and FP32 Jacobi-PCG (FP64 dot
  products) as the inner solver of an FP64 defect correction
  u <- u + K32^-1 (b - K u), where b - K u is the FP64 matrix-free residual
  kernel (never the FP32 matrix), until ||b - K u|| <= max(tol ||b||, atol)
  (Goeddeke, Strzodka, Turek 2007).

Gradients (torch.autograd): the adjoint of feax's f_bwd (linear.py:426-481):
lambda = J_bc^-1 v (J_bc symmetric under symmetric elimination), then
-(dR_bc/dparams)^T lambda from the hand-written VJP kernels.
"""

import numpy as np

from .. import _native as nat
from ..assembler import TracedStructure, assemble_csr, residual
from ..csr import CSRMatrix
from ..DCboundary import DirichletBC
from .common import (KrylovRecord, check_convergence, create_iterative_solve_fn, create_jacobi_preconditioner,
                     create_x0, device_norm, inner_tolerance, krylov, pcg)
from .options import DirectSolverOptions, KrylovSolverOptions, MatrixProperty, resolve_iterative_solver


class SolveStats:
    """What the last solve did (iterations, residuals, outer steps)."""

    def __init__(self):
        self.iters = 0
        self.outer = 0
        self.rs = 0.0
        self.relres = float("nan")
        self.inner_iters = []

    def __repr__(self):
        return (f"SolveStats(iters={self.iters}, outer={self.outer}, relres={self.relres:.3e}, "
                f"inner={self.inner_iters})")


def _b3(options):
    return {"auto": None, "csr": False, "block3": True}[options.spmv]


def solve_fp64(problem, ts, tp, bc, u0, opts, symmetric=True, stats=None, keep=None):
    """One FP64 solve on the assembled CSR (or, operator="matfree", on the
    matrix-free J_bc).  Returns u (FP64 device vector).  keep (dict): receives
    the operator for an adjoint solve."""
    import torch
    if opts.operator == "matfree":
        if problem.ele_type != "HEX8" or problem.vec != 3:
            raise NotImplementedError("the matrix-free operator is the HEX8 (vec = 3) one")
        return _solve_matfree(problem, ts, tp, bc, u0, opts, symmetric, stats, keep)
    if opts.use_jacobi_preconditioner:
        J, dinv = assemble_csr(problem, ts, tp, u0, bc, symmetric, torch.float64, jacobi_shift=opts.jacobi_shift)
    else:
        J, dinv = assemble_csr(problem, ts, tp, u0, bc, symmetric, torch.float64), None
    b = residual(problem, ts, tp, u0, bc, negate=True)
    x0_fn = opts.x0_fn if opts.x0_fn is not None else create_x0(bc.bc_rows, bc.bc_vals)
    x0 = x0_fn(u0)
    du, info = krylov(J, b, x0, opts, dinv, _b3(opts))
    if keep is not None:
        # the adjoint operator J_bc^T is J_bc itself only for a symmetric
        # elimination of a symmetric weak form (not the coupled (u, T) one)
        keep.update(form="fp64", A=J, dinv=dinv, symmetric=symmetric and not getattr(problem, "coupled", False))
    if opts.check_convergence:
        du = check_convergence(J, du, b, opts)
    if stats is not None:
        stats.iters, stats.rs, stats.outer = info.iters, info.rs, 1
        stats.inner_iters = [info.iters]
    u = torch.empty_like(u0)
    nat.call("fx_axpby_f64", nat.stream(), u.numel(), 1.0, nat.ptr(u0), 1.0, nat.ptr(du), nat.ptr(u))
    return u


def solve_gmg(problem, ts, tp, bc, u0, opts, symmetric=True, stats=None, keep=None):
    """FP64 CSR solve preconditioned by a geometric-multigrid V-cycle
    (KrylovSolverOptions.preconditioner = GeometricMultigrid); the V-cycle
    runs in FP32 with precision "mixed", else in FP64."""
    import torch
    from .gmg import GMGHierarchy
    if not symmetric:
        raise NotImplementedError("GeometricMultigrid is built for symmetric elimination")
    J, dinv = assemble_csr(problem, ts, tp, u0, bc, True, torch.float64, jacobi_shift=opts.jacobi_shift)
    b = residual(problem, ts, tp, u0, bc, negate=True)
    x0_fn = opts.x0_fn if opts.x0_fn is not None else create_x0(bc.bc_rows, bc.bc_vals)
    H = GMGHierarchy(problem, ts, tp, u0, bc, J, dinv, opts.preconditioner, fp32=opts.precision != "fp64",
                     jacobi_shift=opts.jacobi_shift)
    du, info = H.solve(b, x0_fn(u0), opts.tol, opts.atol, opts.maxiter)
    if keep is not None:
        keep.update(form="gmg", A=J, dinv=dinv, gmg=H)
    if opts.check_convergence:
        du = check_convergence(J, du, b, opts)
    if stats is not None:
        stats.iters, stats.rs, stats.outer = info.iters, info.rs, 1
        stats.inner_iters = [info.iters]
    u = torch.empty_like(u0)
    nat.call("fx_axpby_f64", nat.stream(), u.numel(), 1.0, nat.ptr(u0), 1.0, nat.ptr(du), nat.ptr(u))
    return u


def _solve_matfree(problem, ts, tp, bc, u0, opts, symmetric, stats, keep):
    """feax's default Krylov path: CG on J_bc v without assembly
    (feax/solvers/linear.py:394-398, feax/assembler.py:1747-1778)."""
    import torch
    from .common import pcg_matfree
    if not symmetric:
        raise NotImplementedError("the matrix-free operator is built for symmetric elimination")
    b = residual(problem, ts, tp, u0, bc, negate=True)
    x0_fn = opts.x0_fn if opts.x0_fn is not None else create_x0(bc.bc_rows, bc.bc_vals)
    du, info = pcg_matfree(problem, ts, tp, None, bc, b, x0_fn(u0), opts.tol, opts.atol, opts.maxiter,
                           opts.check_every)
    if keep is not None:
        keep.update(form="matfree", A=None, dinv=None)
    if stats is not None:
        stats.iters, stats.rs, stats.outer = info.iters, info.rs, 1
        stats.inner_iters = [info.iters]
    u = torch.empty_like(u0)
    nat.call("fx_axpby_f64", nat.stream(), u.numel(), 1.0, nat.ptr(u0), 1.0, nat.ptr(du), nat.ptr(u))
    return u


def solve_mixed(problem, ts, tp, bc, u0, opts, symmetric=True, stats=None, keep=None):
    """FP32 matrix inside an FP64 defect correction.  Returns u (FP64).

    precision "mixed": FP32 difference-form CSR (FX_ASM_DIFF3; in upper-block
    storage, FX_ASM_UPPER, when the law and elimination make it symmetric)
    applied to FP64 CG vectors.  precision "mixed-fp32vec": FP32 CSR and FP32 CG
    vectors (FP64 dot products) on the residual scaled to unit norm.
    The outer residual b - K u is always the FP64 matrix-free kernel.
    """
    import torch
    n = ts.n
    fp32vec = opts.precision == "mixed-fp32vec"
    shift = opts.jacobi_shift if opts.use_jacobi_preconditioner else None
    # The difference form of a linear law with symmetric elimination goes
    # straight into upper-block storage (half the store; its PCG SpMV reads
    # about half the bytes); otherwise the full difference form.
    out = assemble_csr(problem, ts, tp, u0, bc, symmetric, torch.float32, diff_form=not fp32vec, jacobi_shift=shift,
                       upper=not fp32vec)
    A, dinv = out if shift is not None else (out, None)
    if fp32vec and dinv is not None:
        dinv = dinv.to(torch.float32)
    u = u0.clone()
    r = residual(problem, ts, tp, u, bc, negate=True)             # b - K u0 (FP64)
    bnorm = device_norm(r)
    target = max(opts.tol * bnorm, opts.atol)
    vdt = torch.float32 if fp32vec else torch.float64
    rv = torch.empty(n, dtype=vdt, device=u.device) if fp32vec else r
    d = torch.empty(n, dtype=vdt, device=u.device)
    rn = bnorm
    inner, total, outer = [], 0, 0
    budget = opts.maxiter
    # Krylov recycling: the first inner solve records its search directions;
    # later inner solves (same matrix) start from the Galerkin projection of
    # their right-hand side onto them.
    record = None
    if opts.recycle > 0 and not fp32vec:
        record = _record_for(n, min(opts.recycle, opts.maxiter), u.device)
        record.count = 0
        record.owner = object()
    if keep is not None:
        keep.update(form="mixed", A=A, dinv=dinv, record=record,
                    token=record.owner if record is not None else None)
    while rn > target and outer < opts.max_outer and budget > 0:
        if fp32vec:
            nat.call("fx_convert_f64_to_f32", nat.stream(), n, 1.0 / rn, nat.ptr(r), nat.ptr(rv))
        x0 = record.project(rv) if (record is not None and outer > 0) else None
        rec = record if (record is not None and outer == 0) else None
        tin = inner_tolerance(opts, target, rn)
        _, info = pcg(A, rv, x0, tin, 0.0, budget, dinv, True, opts.check_every, out=d, record=rec)
        if fp32vec:
            nat.call("fx_axpy_f32_into_f64", nat.stream(), n, rn, nat.ptr(d), nat.ptr(u))
        else:
            nat.call("fx_axpby_f64", nat.stream(), n, 1.0, nat.ptr(u), 1.0, nat.ptr(d), nat.ptr(u))
        residual(problem, ts, tp, u, bc, negate=True, out=r)
        rn = device_norm(r)
        inner.append(info.iters)
        total += info.iters
        budget -= info.iters
        outer += 1
    if stats is not None:
        stats.iters, stats.outer, stats.inner_iters = total, outer, inner
        stats.rs = rn * rn
        stats.relres = rn / bnorm if bnorm > 0 else 0.0
    if opts.check_convergence and not (rn / (bnorm + 1e-12) < opts.convergence_threshold):
        u = torch.full_like(u, float("nan"))
    return u


_RECORDS = {}


def _record_for(n, capacity, device):
    """Reusable direction store (capacity x n FP32) for Krylov recycling."""
    key = (str(device), n)
    rec = _RECORDS.get(key)
    if rec is None or rec.capacity < capacity:
        _RECORDS.pop(key, None)
        rec = KrylovRecord(n, capacity, device)
        _RECORDS[key] = rec
    return rec


def _dev_vec(a, ts):
    import torch
    a = getattr(a, "dofs", a)
    if isinstance(a, torch.Tensor):
        return a.to(device=ts.device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=ts.device)


def linear_solve(J_bc_applied, res_bc_applied, initial_guess, bc, solver_options, matrix_view=None,
                 traced_params=None, P_mat=None, linear_solve_fn=None, x0_fn=None):
    """One linear step u = u0 + J_bc^-1 (-R_bc(u0)) from user-supplied
    J_bc / R_bc functions (feax/solvers/linear.py:119-160; P_mat, the
    reduced-space prolongation, is not on this path).

    A J_bc in an UPPER or LOWER matrix_view stores one triangle of a
    symmetric matrix; the solve and the convergence check use the full
    matrix A + A^T - diag(A) it stands for (feax/solvers/common.py:364-373;
    feax's direct backends take the view as a factorisation flag)."""
    from ..csr import expand_view
    from .common import create_linear_solve_fn
    if P_mat is not None:
        raise NotImplementedError("prolongation-reduced solves (P_mat) are not on the HIP path")
    if linear_solve_fn is None:
        linear_solve_fn = create_linear_solve_fn(solver_options)
    args = (initial_guess, traced_params) if traced_params is not None else (initial_guess,)
    res = res_bc_applied(*args)
    J = J_bc_applied(*args)
    if isinstance(J, CSRMatrix):
        J = expand_view(J, matrix_view)
    b = -res
    x0 = None
    if solver_options.uses_x0():
        x0 = (x0_fn or create_x0(bc_rows=bc.bc_rows, bc_vals=bc.bc_vals))(initial_guess)
    du = linear_solve_fn(J, b, x0)
    if solver_options.check_convergence:
        du = check_convergence(J, du, b, solver_options, None, "Linear solver")
    return initial_guess + du, None


def create_linear_solver(problem, bc, solver_options=None, adjoint_solver_options=None, traced_params=None,
                         symmetric_elimination=True, traced_structure=None):
    if solver_options is None:
        solver_options = KrylovSolverOptions()
    for o in (solver_options, adjoint_solver_options):
        if o is not None and not isinstance(o, (KrylovSolverOptions, DirectSolverOptions)):
            raise TypeError(f"Expected DirectSolverOptions or KrylovSolverOptions, got {type(o).__name__}.")
    use_direct = isinstance(solver_options, DirectSolverOptions)
    if use_direct and not symmetric_elimination:
        raise NotImplementedError("the direct solver factors the symmetric elimination's SPD J_bc "
                                  "(symmetric_elimination=True)")
    if not use_direct:
        solver_options = resolve_iterative_solver(solver_options, MatrixProperty.SPD)
    if adjoint_solver_options is None:
        adjoint_solver_options = solver_options
    elif isinstance(adjoint_solver_options, KrylovSolverOptions):
        adjoint_solver_options = resolve_iterative_solver(adjoint_solver_options, MatrixProperty.SPD)
    from .gmg import GeometricMultigrid
    use_gmg = not use_direct and isinstance(solver_options.preconditioner, GeometricMultigrid)
    if use_gmg:
        if solver_options.solver != "cg" or solver_options.operator != "csr":
            raise NotImplementedError("GeometricMultigrid preconditions cg on the assembled operator")
    elif not use_direct:
        create_iterative_solve_fn(solver_options)      # validates the options
    default_bc = bc
    ts_default = traced_structure
    stats = SolveStats()

    def forward(tp, ig, eff_bc, ts, keep=None):
        if ts is None:
            ts = getattr(problem, "_default_ts", None) or TracedStructure.from_problem(problem)
            problem._default_ts = ts
        if keep is not None:
            keep["ts"] = ts
        # No initial guess: feax's default, zero_like_initial_guess of the
        # solver's own bc (feax/solvers/linear.py:471-483), also when another
        # bc is passed to the call.
        u0 = _dev_vec(ig, ts) if ig is not None else _zero_like(problem, default_bc, ts)
        if use_direct:
            from .direct import solve_direct
            return solve_direct(problem, ts, tp, eff_bc, u0, solver_options, symmetric_elimination, stats, keep)
        if use_gmg:
            if problem.ele_type != "HEX8" or problem.vec != 3:
                raise NotImplementedError("GeometricMultigrid is built on the structured HEX8 box (vec = 3)")
            return solve_gmg(problem, ts, tp, eff_bc, u0, solver_options, symmetric_elimination, stats, keep)
        if solver_options.solver in ("bicgstab", "gmres") and (solver_options.precision != "fp64"
                                                               or solver_options.operator != "csr"):
            raise NotImplementedError(f"{solver_options.solver} runs on the FP64 CSR operator "
                                      "(precision='fp64', operator='csr')")
        if solver_options.precision != "fp64" and solver_options.operator == "csr":
            if problem.vec != 3:
                raise NotImplementedError("mixed precision uses the block-3 difference form (vec = 3)")
            return solve_mixed(problem, ts, tp, eff_bc, u0, solver_options, symmetric_elimination, stats, keep)
        return solve_fp64(problem, ts, tp, eff_bc, u0, solver_options, symmetric_elimination, stats, keep)

    def solver_wrapper(traced_params, initial_guess=None, bc=None, traced_structure=None):
        eff_bc = bc if isinstance(bc, DirichletBC) else default_bc
        ts = traced_structure if traced_structure is not None else ts_default
        from .adjoint import differentiable_linear_solve
        return differentiable_linear_solve(problem, forward, traced_params, initial_guess, eff_bc, ts,
                                           adjoint_solver_options, symmetric_elimination)

    solver_wrapper.stats = stats
    return solver_wrapper


def _zero_like(problem, bc, ts):
    """zeros with u[bc_rows] = bc_vals (feax/utils.py:464-499), on the device."""
    import torch
    u = torch.empty(ts.n, dtype=torch.float64, device=ts.device)
    nbc = len(bc.bc_rows)
    rows, _, vals = bc.device(ts.device) if nbc else (None, None, None)
    nat.call("fx_bc_lift_f64", nat.stream(), ts.n, nbc, nat.ptr(rows), nat.ptr(vals), None, nat.ptr(u))
    return u
