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
        return np.asarray(eval_points(var_fn, problem.points), float)

    @staticmethod
    def create_cell_var_from_fn(problem, var_fn, var_index=0):
        cent = problem.points[problem.cells].mean(axis=1)
        return np.asarray(eval_points(var_fn, cent), float)
