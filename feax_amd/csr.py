"""CSRMatrix on the device (the feax.csr API).

Mirrors feax/csr.py:23-127: children (data, indptr, indices), static shape,
``nse``, ``dtype``, ``with_data``, ``A @ v``, ``diagonal()``, ``todense()``.
``data`` / ``indptr`` / ``indices`` are device tensors in feax's exact
layout (int32 indptr and indices, row-major, columns sorted); ``@`` and
``diagonal`` run the HIP kernels.

Two extras for HEX8 vec = 3 patterns:
* ``block3``: rows 3m..3m+2 share one column list of whole 3-blocks; the
  SpMV then reads one column index per 3x3 block (``bcol``, built lazily
  on the device by fx_csr_block3_index).
* ``form == "diff3"``: FP32 values in the difference form written by the
  assembler for the mixed-precision solver (the diagonal 3x3 block slot
  holds the block row sums S; off-diagonal blocks act on x_m - x_n).  Its
  ``@`` applies that operator to FP64 vectors; the array is not a plain
  CSR value array.
"""

import numpy as np

from . import _native as nat


class CSRMatrix:
    def __init__(self, data, indptr, indices, shape, block3=False, form="csr", bcol=None):
        self.data = data
        self.indptr = indptr
        self.indices = indices
        self.shape = tuple(shape)
        self.block3 = bool(block3)
        self.form = form
        self._bcol = bcol

    @property
    def nse(self):
        return int(self.indices.shape[0])

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def ndim(self):
        return 2

    def with_data(self, data, form=None):
        return CSRMatrix(data, self.indptr, self.indices, self.shape, self.block3,
                         self.form if form is None else form, self._bcol)

    @property
    def bcol(self):
        """Column node of each 3x3 block (nse/9 int32, device)."""
        import torch
        if not self.block3:
            raise ValueError("bcol needs a block-3 pattern")
        if self._bcol is None:
            b = torch.empty(self.nse // 9, dtype=torch.int32, device=self.data.device)
            nat.call("fx_csr_block3_index", nat.stream(), self.shape[0], nat.ptr(self.indptr),
                     nat.ptr(self.indices), nat.ptr(b))
            self._bcol = b
        return self._bcol

    def _sfx(self):
        import torch
        if self.data.dtype == torch.float64:
            return "_f64"
        if self.data.dtype == torch.float32:
            return "_f32"
        raise TypeError(f"CSR data dtype {self.data.dtype} not supported")

    def matvec(self, v, out=None, block3=None):
        """y = A v with the HIP SpMV (scalar CSR, block-3, or difference form)."""
        import torch
        n = self.shape[0]
        if v.shape[0] != self.shape[1]:
            raise ValueError(f"matvec: vector length {v.shape[0]} != {self.shape[1]}")
        if self.form == "diff3":
            v = v.to(torch.float64).contiguous()
            y = torch.empty(n, dtype=torch.float64, device=self.data.device) if out is None else out
            nat.call("fx_csr_spmv_diff3_f32f64", nat.stream(), n, self.nse, nat.ptr(self.indptr),
                     nat.ptr(self.bcol), nat.ptr(self.data), nat.ptr(v), nat.ptr(y))
            return y
        v = v.to(self.data.dtype).contiguous()
        y = torch.empty(n, dtype=self.data.dtype, device=self.data.device) if out is None else out
        use_b3 = self.block3 if block3 is None else (block3 and self.block3)
        if use_b3:
            nat.call("fx_csr_spmv_b3" + self._sfx(), nat.stream(), n, self.nse, nat.ptr(self.indptr),
                     nat.ptr(self.bcol), nat.ptr(self.data), nat.ptr(v), nat.ptr(y))
        else:
            nat.call("fx_csr_spmv" + self._sfx(), nat.stream(), n, self.nse, nat.ptr(self.indptr),
                     nat.ptr(self.indices), nat.ptr(self.data), nat.ptr(v), nat.ptr(y))
        return y

    def __matmul__(self, v):
        return self.matvec(v)

    def diagonal(self):
        import torch
        if self.form != "csr":
            raise ValueError("diagonal() of a difference-form matrix is not its diagonal")
        n = self.shape[0]
        d = torch.empty(n, dtype=self.data.dtype, device=self.data.device)
        nat.call("fx_csr_diagonal" + self._sfx(), nat.stream(), n, nat.ptr(self.indptr), nat.ptr(self.indices),
                 nat.ptr(self.data), nat.ptr(d))
        return d

    def todense(self):
        """Dense host copy (tests / tiny problems)."""
        if self.form != "csr":
            raise ValueError("todense() of a difference-form matrix")
        data = self.data.detach().cpu().numpy()
        indptr = self.indptr.cpu().numpy()
        indices = self.indices.cpu().numpy()
        rows = np.repeat(np.arange(self.shape[0]), np.diff(indptr))
        out = np.zeros(self.shape, dtype=data.dtype)
        np.add.at(out, (rows, indices), data)
        return out

    def to_scipy(self):
        import scipy.sparse as sp
        if self.form != "csr":
            raise ValueError("to_scipy() of a difference-form matrix")
        return sp.csr_matrix((self.data.detach().cpu().numpy(), self.indices.cpu().numpy(),
                              self.indptr.cpu().numpy()), shape=self.shape)


def transpose_with_maps(J, T_perm, T_indptr, T_indices):
    """Explicit CSR transpose from the static maps of Problem
    (csr_T_perm / csr_T_indptr / csr_T_indices): values gathered as
    J.data[T_perm] (feax/csr.py:125-141)."""
    import torch
    dev = J.data.device
    perm = torch.as_tensor(np.asarray(T_perm, np.int64), device=dev)
    ip = torch.as_tensor(np.asarray(T_indptr, np.int32), device=dev)
    ix = torch.as_tensor(np.asarray(T_indices, np.int32), device=dev)
    return CSRMatrix(J.data[perm], ip, ix, (J.shape[1], J.shape[0]), block3=J.block3)
