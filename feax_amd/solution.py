"""Solution: flat DOF vector + per-variable layout (feax/solution.py:42-173)."""

import numpy as np


class Solution:
    def __init__(self, dofs, layout):
        self.dofs = dofs
        self.layout = tuple((int(n), int(v)) for n, v in layout)

    @classmethod
    def from_problem(cls, problem, dofs):
        return cls(dofs, problem.solution_layout)

    def __array__(self, dtype=None, copy=None):
        d = self.dofs
        a = d.detach().cpu().numpy() if hasattr(d, "detach") else np.asarray(d)
        return a.astype(dtype) if dtype is not None else a

    @property
    def flat(self):
        return self.dofs

    @property
    def shape(self):
        return tuple(self.dofs.shape)

    @property
    def dtype(self):
        return self.dofs.dtype

    @property
    def num_fields(self):
        return len(self.layout)

    def _other(self, o):
        return o.dofs if isinstance(o, Solution) else o

    def __add__(self, o):
        return self.dofs + self._other(o)

    def __radd__(self, o):
        return self._other(o) + self.dofs

    def __sub__(self, o):
        return self.dofs - self._other(o)

    def __rsub__(self, o):
        return self._other(o) - self.dofs

    def __mul__(self, o):
        return self.dofs * self._other(o)

    def __rmul__(self, o):
        return self._other(o) * self.dofs

    def __truediv__(self, o):
        return self.dofs / self._other(o)

    def __rtruediv__(self, o):
        return self._other(o) / self.dofs

    def __pow__(self, o):
        return self.dofs ** self._other(o)

    def __neg__(self):
        return -self.dofs

    def __matmul__(self, o):
        return self.dofs @ self._other(o)

    def __rmatmul__(self, o):
        return self._other(o) @ self.dofs

    def block_until_ready(self):
        """Waits for the device work that produces the DOFs."""
        if hasattr(self.dofs, "is_cuda") and self.dofs.is_cuda:
            import torch
            torch.cuda.synchronize(self.dofs.device)
        return self

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        """torch functions see the flat DOF vector (the role of feax's
        __jax_array__): torch.sum(sol), torch.dot(sol, v), ..."""
        def unwrap(a):
            if isinstance(a, Solution):
                return a.dofs
            if isinstance(a, (list, tuple)):
                return type(a)(unwrap(x) for x in a)
            return a
        return func(*unwrap(args), **{k: unwrap(v) for k, v in (kwargs or {}).items()})

    def __getitem__(self, key):
        return self.dofs[key]

    def field(self, var_index=0):
        if not isinstance(var_index, (int, np.integer)):
            raise TypeError(f"field() takes an int (got {type(var_index).__name__})")
        sizes = [n * v for n, v in self.layout]
        if not 0 <= var_index < len(sizes):
            raise IndexError(f"var_index {var_index} out of range ({len(sizes)} field(s))")
        start = sum(sizes[:var_index])
        n, v = self.layout[var_index]
        seg = self.dofs[..., start:start + n * v]
        return seg.reshape(*seg.shape[:-1], n, v)

    def node_var(self, component=None, var_index=0):
        u = self.field(var_index)
        vec = self.layout[var_index][1]
        if vec == 1:
            return u[..., 0]
        if component is None:
            raise ValueError(f"field has vec={vec} components; pass component=")
        return u[..., component]

    def __repr__(self):
        fields = ", ".join(f"({n}, {v})" for n, v in self.layout)
        return f"Solution(dofs={tuple(self.dofs.shape)}, fields=[{fields}])"
