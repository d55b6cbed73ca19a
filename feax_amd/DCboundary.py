"""Dirichlet boundary conditions (the feax.DCboundary API).

Mirrors feax/DCboundary.py: DirichletBC (bc_rows, bc_mask, bc_vals,
total_dofs; replace_vals, from_specs), DirichletBCSpec, DirichletBCConfig
(add, create_bc), dirichlet_bc_config.  Rows are concatenated per spec and
component, sorted, and duplicates keep their first occurrence
(feax/DCboundary.py:395-501).  The BC is applied on the device by the
element kernels (CSR elimination fused into the assembly store, residual
rows set in the residual kernel).
"""

from dataclasses import dataclass, field
from typing import Callable, List, Union

import numpy as np

from .fe import eval_points, location_flags


@dataclass(frozen=True)
class DirichletBC:
    bc_rows: np.ndarray
    bc_mask: np.ndarray
    bc_vals: np.ndarray
    total_dofs: int

    def replace_vals(self, new_bc_vals):
        """Same DOFs, new prescribed values (feax/DCboundary.py:36-61).  A torch
        tensor is kept as it is, so a solve can differentiate with respect to
        it (the bc_vals cotangent of feax's custom VJPs)."""
        if type(new_bc_vals).__module__.startswith("torch"):
            if tuple(new_bc_vals.shape) != (len(self.bc_rows),):
                raise ValueError(f"bc_vals of shape {tuple(new_bc_vals.shape)} for {len(self.bc_rows)} rows")
            return DirichletBC(self.bc_rows, self.bc_mask, new_bc_vals, self.total_dofs)
        return DirichletBC(self.bc_rows, self.bc_mask, np.asarray(new_bc_vals, np.float64), self.total_dofs)

    def vals_numpy(self):
        """bc_vals as a float64 numpy array (a tensor is detached and copied)."""
        v = self.bc_vals
        if type(v).__module__.startswith("torch"):
            return v.detach().to("cpu", dtype=__import__("torch").float64).numpy()
        return np.asarray(v, np.float64)

    @staticmethod
    def from_specs(problem, specs):
        return DirichletBCConfig(list(specs)).create_bc(problem)

    def device(self, dev="cuda"):
        """Device copies (bc_rows int32, bc_mask uint8, bc_vals float64)."""
        import torch
        key = str(dev)
        cache = self.__dict__.setdefault("_dev", {})
        if key not in cache:
            vals = self.bc_vals
            vals = vals.detach().to(device=dev, dtype=torch.float64).contiguous() if isinstance(vals, torch.Tensor) \
                else torch.as_tensor(np.ascontiguousarray(vals, np.float64), device=dev)
            cache[key] = (torch.as_tensor(np.ascontiguousarray(self.bc_rows, np.int32), device=dev),
                          torch.as_tensor(np.ascontiguousarray(self.bc_mask, np.uint8), device=dev), vals)
        return cache[key]


@dataclass
class DirichletBCSpec:
    """location(point) -> bool, component int | 'x' | 'y' | 'z' | 'all',
    value float | fn(point) (feax/DCboundary.py:259-321)."""
    location: Callable
    component: Union[int, str]
    value: Union[float, Callable]
    variable_index: Union[int, None] = None

    def __post_init__(self):
        if isinstance(self.component, str):
            cmap = {"x": 0, "y": 1, "z": 2, "all": "all"}
            if self.component.lower() not in cmap:
                raise ValueError(f"Invalid component string: {self.component}. Must be 'x', 'y', 'z', or 'all'")
            self.component = cmap[self.component.lower()]
        elif isinstance(self.component, (int, np.integer)):
            if self.component < 0:
                raise ValueError(f"Component index must be non-negative, got {self.component}")
        if isinstance(self.value, (int, float)):
            const = float(self.value)
            self.value = lambda pt, _c=const: _c


@dataclass
class DirichletBCConfig:
    specs: List[DirichletBCSpec] = field(default_factory=list)

    def add(self, location, component, value):
        self.specs.append(DirichletBCSpec(location, component, value))
        return self

    def create_bc(self, problem):
        n = problem.num_total_dofs_all_vars
        if not self.specs:
            return DirichletBC(np.zeros(0, np.int32), np.zeros(n, bool), np.zeros(0, np.float64), n)
        rows, vals = [], []
        # per variable and spec, as feax/DCboundary.py:427-480: the spec's
        # component(s) of the variables it names (variable_index None: all)
        for ind, fe in enumerate(problem.fes):
            vec, off = fe.vec, problem.offset[ind]
            pts = np.asarray(fe.points, np.float64)
            for spec in self.specs:
                if spec.variable_index is not None and spec.variable_index != ind:
                    continue
                comps = list(range(vec)) if spec.component == "all" else (
                    [spec.component] if spec.component < vec else [])
                if not comps:
                    continue
                # The location is evaluated once per spec (feax evaluates it per
                # component, with the same result).
                nodes = np.nonzero(location_flags(spec.location, pts, "got"))[0]
                if len(nodes) == 0:
                    continue
                v = np.asarray(eval_points(spec.value, pts[nodes]), np.float64).reshape(-1)
                if v.shape != (len(nodes),):
                    raise ValueError(f"a Dirichlet value function must return one value per point "
                                     f"(got {v.size} values for {len(nodes)} points)")
                for c in comps:
                    rows.append(nodes * vec + c + off)
                    vals.append(v)
        if rows:
            r = np.concatenate(rows)
            v = np.concatenate(vals)
            s = np.argsort(r, kind="stable")
            r, v = r[s], v[s]
            r, ui = np.unique(r, return_index=True)
            v = v[ui]
        else:
            r, v = np.zeros(0, np.int64), np.zeros(0)
        mask = np.zeros(n, bool)
        mask[r] = True
        return DirichletBC(r.astype(np.int32), mask, v.astype(np.float64), n)


def dirichlet_bc_config(*specs):
    return DirichletBCConfig(list(specs))
