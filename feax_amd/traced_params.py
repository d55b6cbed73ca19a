"""TracedParams: the differentiable inputs of a solve (the feax API).

Mirrors feax/traced_params.py:41-375: ``volume_vars`` (tuple; node (N,),
cell (E,) or quad (E, 8) arrays) and ``surface_vars`` (list, one tuple per
location_fn, arrays of shape (F_i, 4)), with the create_* helpers.  Arrays
may be numpy or torch; the solver moves them to the device.
"""

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .fe import eval_points


@dataclass(frozen=True)
class TracedParams:
    volume_vars: Tuple = ()
    surface_vars: Optional[List[Tuple]] = None

    def __post_init__(self):
        if self.surface_vars is None:
            object.__setattr__(self, "surface_vars", [])
        vv = tuple(getattr(v, "dofs", v) for v in self.volume_vars)
        object.__setattr__(self, "volume_vars", vv)

    @staticmethod
    def node_var_from_solution(problem, sol, component=None, var_index=0):
        """A solved field as a (num_nodes,) node variable of the next solve
        on the same mesh (feax/traced_params.py:378-423); a torch solution
        stays differentiable.  component is required when vec > 1."""
        sol = getattr(sol, "dofs", sol)
        if var_index != 0:
            raise IndexError("single-variable problems have var_index 0 only")
        u = sol.reshape(problem.num_total_nodes, problem.vec) if hasattr(sol, "reshape") else \
            np.asarray(sol).reshape(problem.num_total_nodes, problem.vec)
        if problem.vec == 1:
            return u[:, 0]
        if component is None:
            raise ValueError(f"solution has vec={problem.vec} components; pass component=")
        return u[:, component]

    def replace_volume_var(self, index, new_var):
        """A copy with volume_vars[index] replaced (feax/traced_params.py:425-442)."""
        vv = list(self.volume_vars)
        vv[index] = new_var
        return TracedParams(volume_vars=tuple(vv), surface_vars=list(self.surface_vars))

    def replace_surface_var(self, surface_index, var_index, new_var):
        """A copy with surface_vars[surface_index][var_index] replaced
        (feax/traced_params.py:444-466)."""
        sv = [list(t) for t in self.surface_vars]
        sv[surface_index][var_index] = new_var
        return TracedParams(volume_vars=self.volume_vars, surface_vars=[tuple(t) for t in sv])

    @staticmethod
    def create_node_var(problem, value, var_index=0):
        return np.full(problem.fes[var_index].num_total_nodes, float(value))

    @staticmethod
    def create_cell_var(problem, value, var_index=0):
        return np.full(problem.num_cells, float(value))

    @staticmethod
    def create_uniform_volume_var(problem, value, var_index=0):
        return np.full((problem.num_cells, problem.fes[var_index].num_quads), float(value))

    @staticmethod
    def create_uniform_surface_var(problem, value, surface_index=0):
        nf = len(problem.boundary_inds_list[surface_index])
        return np.full((nf, problem.fes[0].num_face_quads), float(value))

    @staticmethod
    def create_node_var_from_fn(problem, var_fn, var_index=0):
        """var_fn at every node (feax/traced_params.py:234-266)."""
        return np.asarray(eval_points(var_fn, problem.fes[var_index].points), float)

    @staticmethod
    def create_cell_var_from_fn(problem, var_fn, var_index=0):
        """var_fn at every cell centroid, the mean of its vertices
        (feax/traced_params.py:269-302)."""
        fe = problem.fes[var_index]
        cent = np.asarray(fe.points, float)[np.asarray(fe.cells)].mean(axis=1)
        return np.asarray(eval_points(var_fn, cent), float)

    @staticmethod
    def create_spatially_varying_volume_var(problem, var_fn, var_index=0):
        """var_fn at every volume quadrature point, shape (E, num_quads)
        (feax/traced_params.py:304-340)."""
        qp = problem.physical_quad_points
        E, nq, dim = qp.shape
        return np.asarray(eval_points(var_fn, qp.reshape(-1, dim)), float).reshape(E, nq)

    @staticmethod
    def create_spatially_varying_surface_var(problem, var_fn, surface_index=0):
        """var_fn at every face quadrature point of boundary surface_index,
        shape (F, num_face_quads) (feax/traced_params.py:342-375)."""
        sp = problem.physical_surface_quad_points[surface_index]
        F, nfq, dim = sp.shape
        return np.asarray(eval_points(var_fn, sp.reshape(-1, dim)), float).reshape(F, nfq)
