"""create_solver: the drop-in entry point (feax/solver.py:48-416).

Routes ``linear=True`` to create_linear_solver and ``linear=False`` to the
Newton solver; wraps the flat result in a Solution when
``return_solution=True`` (feax/solver.py:232-246).
"""

from .solution import Solution
from .solvers.linear import create_linear_solver
from .solvers.options import DirectSolverOptions


def create_solver(problem, bc, solver_options=None, adjoint_solver_options=None, newton_options=None,
                  linear=False, P=None, traced_params=None, extra_residual_fn=None, symmetric_elimination=True,
                  traced_structure=None, return_solution=True):
    if P is not None:
        raise NotImplementedError("periodic (P) reduced solves are not on the HIP path")
    if extra_residual_fn is not None:
        raise NotImplementedError("extra_residual_fn is not on the HIP path")
    if solver_options is None:
        # feax's default (feax/solver.py:293-301): a direct solver.  Its
        # "auto" needs no sample traced_params here (the factorisation picks
        # Cholesky or L D L^T by itself), so feax's ValueError for a missing
        # traced_params does not arise.
        solver_options = DirectSolverOptions()

    def _wrap(fn):
        if not return_solution:
            return fn
        layout = ((problem.num_total_nodes, problem.vec),)

        def wrapped(*args, **kwargs):
            args = tuple(a.dofs if isinstance(a, Solution) else a for a in args)
            kwargs = {k: (v.dofs if isinstance(v, Solution) else v) for k, v in kwargs.items()}
            return Solution(fn(*args, **kwargs), layout)
        wrapped.stats = getattr(fn, "stats", None)
        return wrapped

    if linear:
        return _wrap(create_linear_solver(problem, bc, solver_options, adjoint_solver_options, traced_params,
                                          symmetric_elimination, traced_structure))
    from .solvers.newton import create_newton_solver
    return _wrap(create_newton_solver(problem, bc, solver_options, adjoint_solver_options, newton_options,
                                      traced_params, symmetric_elimination, traced_structure))
