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
  holds the block row sums S; off-diagonal blocks act on x_m - x_n), stored
  block by block: a node's values start where its three CSR rows start
  (indptr[3 m]) and hold its 3x3 blocks one after another, nine values each,
  row-major (one contiguous run per block for the SpMV lane that reads it).
  Its ``@`` applies that operator to FP64 vectors; the array is not a plain
  CSR value array.
* ``form == "sym3"``: the difference form of a bitwise-symmetric matrix in
  upper-block storage (``data`` holds node n's blocks with column >= n, nine
  FP32 values each, from block ``up_ptr[n]`` of the pattern's fx_b3sym
  tables; a lower block is the transpose of its mirror): the same operator
  as "diff3" from about half the values, assembled directly by the HEX8
  FP32 assembly (FX_ASM_UPPER).
"""

import ctypes

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
        # callable(value_bytes) -> block-column map of the pattern (or None),
        # set by the assembly for block-3 matrices; used by the PCG SpMV
        self.b3map = None
        # callable() -> upper-block storage map of the pattern (or None), set
        # by the assembly for block-3 matrices (see upper_storage)
        self.b3sym = None
        self._upper = None
        # set by the assembly when the matrix is symmetric bit for bit by
        # construction (upper_storage is then worth its compaction pass)
        self.sym_hint = False

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
        A = CSRMatrix(data, self.indptr, self.indices, self.shape, self.block3,
                      self.form if form is None else form, self._bcol)
        A.b3map, A.b3sym = self.b3map, self.b3sym
        A.sym_hint = self.sym_hint and form in (None, self.form)
        return A

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

    def upper_storage(self):
        """(nat.B3Sym, upper values) of a bitwise-symmetric block-3 matrix,
        else None: the FP32 difference form (form "diff3"; "sym3" is already
        stored so) or the FP64 block-3 CSR.  The upper 3x3 blocks are copied
        once into the compact storage of fx_b3sym_compact_* (about half the
        values), which also checks every lower block against the transpose of
        its mirror; the PCG then runs its SpMV on the compact copy
        (fx_pcg_b3sym_*): the same operator from about half the bytes.  Only
        tried for matrices the assembly marks symmetric by construction
        (sym_hint: a linear law with symmetric elimination); one that is not
        exactly symmetric keeps the full storage.  Built once per matrix (one
        host sync)."""
        import torch
        if self.form == "sym3":
            return self.b3sym(64), self.data
        fp64 = self.form == "csr" and self.block3 and self.data.dtype == torch.float64
        if not (self.form == "diff3" or fp64) or self.b3sym is None or not self.sym_hint:
            return None
        if self._upper is None:
            self._upper = False
            sym = self.b3sym(32 if fp64 else 64)
            if sym is not None:
                dev = self.data.device
                upper = torch.empty(9 * sym.num_upper, dtype=self.data.dtype, device=dev)
                bad = torch.empty(2, dtype=torch.int32, device=dev)
                nat.call("fx_b3sym_compact_f64" if fp64 else "fx_b3sym_compact_f32", nat.stream(), self.shape[0],
                         nat.ptr(self.indptr), nat.ptr(self.bcol), ctypes.addressof(sym.struct), nat.ptr(self.data),
                         nat.ptr(upper), nat.ptr(bad))
                if int(bad.abs().sum().item()) == 0:
                    self._upper = (sym, upper)
        return self._upper or None

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
        if self.form == "sym3":
            v = v.to(torch.float64).contiguous()
            y = torch.empty(n, dtype=torch.float64, device=self.data.device) if out is None else out
            sym = self.b3sym(64)
            nat.call("fx_csr_spmv_b3sym_f32f64", nat.stream(), n, nat.ptr(self.indptr), ctypes.addressof(sym.struct),
                     nat.ptr(self.data), nat.ptr(v), nat.ptr(y))
            return y
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

    def transpose(self):
        """The explicit transpose as a CSRMatrix (feax/csr.py:102-113 gives a
        matvec-only view; here the transposed values are one device gather
        of ``data`` through a slot map built once per pattern on the host)."""
        if self.form != "csr":
            raise ValueError("transpose() of a difference-form matrix")
        ip, ix, perm = _transpose_maps(self)
        return CSRMatrix(self.data[perm], ip, ix, (self.shape[1], self.shape[0]), block3=self.block3)

    @property
    def T(self):
        return self.transpose()

    def rmatvec(self, w):
        """A^T w (feax/csr.py:92-100)."""
        return self.transpose().matvec(w)

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


_T_CACHE = {}


def _transpose_maps(A):
    """(indptr, indices, perm) of A's transposed pattern, cached per pattern:
    slot k of the transpose holds A.data[perm[k]], rows sorted, columns
    sorted within a row (feax/problem.py:417-451)."""
    import torch
    key = (A.indptr.data_ptr(), A.indices.data_ptr(), A.shape)
    hit = _T_CACHE.get(key)
    if hit is None:
        n, m = A.shape
        indptr = A.indptr.cpu().numpy().astype(np.int64)
        cols = A.indices.cpu().numpy().astype(np.int64)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
        perm = np.argsort(cols * n + rows, kind="stable")
        t_indptr = np.zeros(m + 1, np.int64)
        np.cumsum(np.bincount(cols, minlength=m), out=t_indptr[1:])
        dev = A.data.device
        hit = (torch.as_tensor(t_indptr.astype(np.int32), device=dev),
               torch.as_tensor(rows[perm].astype(np.int32), device=dev),
               torch.as_tensor(perm, device=dev), A.indptr, A.indices)
        _T_CACHE[key] = hit
    return hit[:3]


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


_VIEW_CACHE = {}


def expand_view(A, matrix_view):
    """The full symmetric matrix a triangular view stands for: A + A^T -
    diag(A) (feax/solvers/common.py:364-373, _matvec_with_matrix_view).

    ``matrix_view`` is a MatrixView (or its name); FULL and None return A.
    The full pattern and the gather map from the triangle's slots are built
    once per pattern on the host (rows sorted, columns sorted within a row,
    as feax's CSR); the values are one device gather per call.
    """
    import torch
    name = getattr(matrix_view, "name", matrix_view)
    if name in (None, "FULL"):
        return A
    if name not in ("UPPER", "LOWER"):
        raise ValueError(f"unknown matrix_view {matrix_view!r}")
    if A.form != "csr":
        raise ValueError("expand_view needs a plain CSR matrix")
    key = (A.indptr.data_ptr(), A.indices.data_ptr(), A.shape, name)
    hit = _VIEW_CACHE.get(key)
    if hit is None:
        n = A.shape[0]
        indptr = A.indptr.cpu().numpy().astype(np.int64)
        cols = A.indices.cpu().numpy().astype(np.int64)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
        keep = cols >= rows if name == "UPPER" else cols <= rows
        if not np.all(keep):
            raise ValueError(f"matrix_view {name}: the matrix has entries outside that triangle")
        off = cols != rows
        slot = np.arange(len(cols), dtype=np.int64)
        r_all = np.concatenate([rows, cols[off]])
        c_all = np.concatenate([cols, rows[off]])
        s_all = np.concatenate([slot, slot[off]])
        order = np.argsort(r_all * n + c_all, kind="stable")
        full_indptr = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(r_all, minlength=n), out=full_indptr[1:])
        dev = A.data.device
        hit = (torch.as_tensor(full_indptr.astype(np.int32), device=dev),
               torch.as_tensor(c_all[order].astype(np.int32), device=dev),
               torch.as_tensor(s_all[order], device=dev),
               A.indptr, A.indices)          # keep the keyed tensors alive
        _VIEW_CACHE[key] = hit
    ip, ix, src = hit[:3]
    return CSRMatrix(A.data[src], ip, ix, A.shape)
