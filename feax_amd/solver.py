"""create_solver: the drop-in entry point (feax/solver.py:48-416).

Routes ``linear=True`` to create_linear_solver and ``linear=False`` to the
Newton solver; wraps the flat result in a Solution when
``return_solution=True`` (feax/solver.py:232-246).  Missing options default
to DirectSolverOptions(); a "auto" solver is resolved from a sample J_bc
assembled at zero_like_initial_guess with the given traced_params
(feax/solver.py:293-373), which is then required.
"""

from .solution import Solution
from .solvers.linear import create_linear_solver
from .solvers.options import (DirectSolverOptions, KrylovSolverOptions, detect_matrix_property, resolve_direct_solver,
                              resolve_iterative_solver)


def _is_resolvable_auto(o):
    return isinstance(o, (DirectSolverOptions, KrylovSolverOptions)) and o.solver == "auto"


def _resolve(o, mp, matrix_view, role):
    cat = "direct" if isinstance(o, DirectSolverOptions) else "iterative"
    r = resolve_direct_solver(o, mp, matrix_view) if cat == "direct" else resolve_iterative_solver(o, mp)
    print(f"[feax] Auto {role} ({cat}): backend=HIP, matrix_property={mp.name} -> {r.solver}")
    return r


def sample_matrix_property(problem, bc, traced_params, symmetric_elimination=True, traced_structure=None):
    """MatrixProperty of J_bc at zero_like_initial_guess (feax/solver.py:333-352):
    one FP64 assembly with the given elimination, then detect_matrix_property."""
    import torch
    from .assembler import TracedStructure, assemble_csr, view_csr
    from .solvers.linear import _zero_like
    ts = traced_structure
    if ts is None:
        ts = getattr(problem, "_default_ts", None) or TracedStructure.from_problem(problem)
        problem._default_ts = ts
    u0 = _zero_like(problem, bc, ts)
    J = view_csr(problem, ts, assemble_csr(problem, ts, traced_params, u0, bc, symmetric_elimination, torch.float64))
    return detect_matrix_property(J, matrix_view=problem.matrix_view)


def create_solver(problem, bc, solver_options=None, adjoint_solver_options=None, newton_options=None,
                  linear=False, P=None, traced_params=None, extra_residual_fn=None, symmetric_elimination=True,
                  traced_structure=None, return_solution=True):
    if P is not None:
        raise NotImplementedError("periodic (P) reduced solves are not on the HIP path")
    if extra_residual_fn is not None:
        raise NotImplementedError("extra_residual_fn is not on the HIP path")
    lin, adj = solver_options, adjoint_solver_options
    missing = lin is None
    if lin is None:
        # feax's default (feax/solver.py:293-301): a direct solver, "auto".
        lin = DirectSolverOptions()
    if adj is None:
        adj = lin
    shared = adj is lin
    needs_auto = missing or _is_resolvable_auto(lin) or _is_resolvable_auto(adj)
    if needs_auto and traced_params is None:
        raise ValueError(
            "traced_params is required when solver_options is None or has solver='auto'. "
            "Pass a sample TracedParams to enable automatic matrix property detection, "
            "or specify the solver explicitly (e.g. KrylovSolverOptions(solver='cg')).")
    for role, o in (("solver_options", lin), ("adjoint_solver_options", adj)):
        if not isinstance(o, (DirectSolverOptions, KrylovSolverOptions)):
            raise TypeError(f"Unsupported {role} type. Expected DirectSolverOptions or KrylovSolverOptions, "
                            f"got {type(o).__name__}.")
    if needs_auto:
        mp = sample_matrix_property(problem, bc, traced_params, symmetric_elimination, traced_structure)
        if _is_resolvable_auto(lin):
            lin = _resolve(lin, mp, problem.matrix_view, "solver")
        if shared:
            adj = lin
        elif _is_resolvable_auto(adj):
            adj = _resolve(adj, mp, problem.matrix_view, "adjoint solver")

    def _wrap(fn):
        if not return_solution:
            return fn
        layout = problem.solution_layout

        def wrapped(*args, **kwargs):
            args = tuple(a.dofs if isinstance(a, Solution) else a for a in args)
            kwargs = {k: (v.dofs if isinstance(v, Solution) else v) for k, v in kwargs.items()}
            return Solution(fn(*args, **kwargs), layout)
        wrapped.stats = getattr(fn, "stats", None)
        return wrapped

    if linear:
        # feax's linear path assembles with the symmetric elimination whatever
        # symmetric_elimination says (feax/solver.py:394-403).
        return _wrap(create_linear_solver(problem, bc, lin, adj, traced_params, traced_structure=traced_structure))
    from .solvers.newton import create_newton_solver
    return _wrap(create_newton_solver(problem, bc, lin, adj, newton_options, traced_params, symmetric_elimination,
                                      traced_structure))
