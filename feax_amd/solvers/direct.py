"""Direct linear solve on the device: banded Cholesky (or L D L^T) of J_bc.

The stand-in for feax's GPU direct solver (DirectSolverOptions "auto" /
"cudss", feax/solvers/common.py:116-210).  feax's direct branch of
create_linear_solver (feax/solvers/linear.py:308-417) computes
u = u0 + J_bc^-1 (-R_bc(u0)) with no initial iterate, and under a symmetric
operator reuses the forward factorisation for the adjoint solve
(linear.py:351-378); this module does the same with
fx_band_from_csr_f64 / fx_band_cholesky_f64 / fx_band_solve_f64.  A
symmetric matrix with a non-positive pivot (a Neo-Hookean tangent away from
equilibrium) is factored again as L D L^T without pivoting
(fx_band_ldlt_f64), as a symmetric-indefinite direct solver would.

Ordering: the DOFs are renumbered node-wise (the vec DOFs of a node stay
together) by whichever of the mesh's own numbering and reverse
Cuthill-McKee of the node graph gives the smaller bandwidth bw; the band
holds n (bw + 256) FP64 values (44 GB at the 200x40x40 cantilever).
"""

import ctypes

import numpy as np

from .. import _native as nat


def _bandwidth(rows, cols, inv):
    if len(cols) == 0:
        return 0
    if inv is not None:
        rows = inv[rows]
        cols = inv[cols]
    return int(np.abs(rows.astype(np.int64) - cols.astype(np.int64)).max())


def band_ordering(problem):
    """(perm, inv, bw): DOF renumbering new -> old and old -> new (None for
    the identity) with its bandwidth; cached on the problem.  Reverse
    Cuthill-McKee runs on the node graph (each node's DOFs stay together)
    and is kept only when it narrows the band."""
    cached = getattr(problem, "_band_ordering", None)
    if cached is not None:
        return cached
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    indptr = np.asarray(problem.csr_indptr)
    cols = np.asarray(problem.csr_indices)
    n = len(indptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(indptr))
    best = (None, None, _bandwidth(rows, cols, None))
    v = int(getattr(problem, "vec", 3))
    if n % v == 0 and n > v:
        sel = (rows % v == 0) & (cols % v == 0)
        N = n // v
        G = sp.csr_matrix((np.ones(int(sel.sum()), np.int8), (rows[sel] // v, cols[sel] // v)), shape=(N, N))
        node_perm = reverse_cuthill_mckee(G, symmetric_mode=True).astype(np.int64)
        perm = (v * node_perm[:, None] + np.arange(v)).reshape(-1)
        inv = np.empty(n, np.int64)
        inv[perm] = np.arange(n)
        bw = _bandwidth(rows, cols, inv)
        if bw < best[2]:
            best = (perm.astype(np.int32), inv.astype(np.int32), bw)
    problem._band_ordering = best
    return best


class BandCholesky:
    """L L^T = P J P^T of an FP64 CSRMatrix on the device, and solves with it.
    The band storage is reused across factorisations of one size; `owner`
    names the factorisation that last wrote it."""

    _buffers = {}

    def __init__(self, problem, device):
        import torch
        if getattr(problem, "coupled", False):
            raise NotImplementedError("the banded direct solver factors symmetric J_bc; the coupled (u, T) "
                                      "Jacobian is not symmetric (use KrylovSolverOptions(solver='gmres' or "
                                      "'bicgstab'))")
        perm, inv, bw = band_ordering(problem)
        self.n = len(np.asarray(problem.csr_indptr)) - 1
        self.bw = bw
        self.ldab = int(nat.lib().fx_band_ldab(bw))
        self.device = device
        self.perm = None if perm is None else torch.as_tensor(perm, device=device)
        self.inv = None if inv is None else torch.as_tensor(inv, device=device)
        self.ws = torch.empty(int(nat.lib().fx_band_workspace_bytes(self.n, bw)), dtype=torch.uint8, device=device)
        key = (str(device), self.n * self.ldab)
        ab = BandCholesky._buffers.get(key)
        if ab is None:
            ab = torch.empty(self.n * self.ldab, dtype=torch.float64, device=device)
            BandCholesky._buffers[key] = ab
        self.ab = ab
        # inverses of the diagonal blocks (fx_band_invert_diag_f64), formed
        # after each factorisation, and the workspace of the solves that use
        # them (fx_band_solve_inv_f64)
        ikey = (str(device), "inv", int(nat.lib().fx_band_inv_bytes(self.n, bw)))
        inv_blocks = BandCholesky._buffers.get(ikey)
        if inv_blocks is None:
            inv_blocks = torch.empty(ikey[2] // 8, dtype=torch.float64, device=device)
            BandCholesky._buffers[ikey] = inv_blocks
        self.inv_blocks = inv_blocks
        # kept per size like the band storage: the library caches the graph of
        # the blocked solves per (band, inverse blocks, workspace) addresses
        wkey = (str(device), "solve_ws", int(nat.lib().fx_band_solve_inv_workspace_bytes(self.n, bw)))
        ws_solve = BandCholesky._buffers.get(wkey)
        if ws_solve is None:
            ws_solve = torch.empty(wkey[2], dtype=torch.uint8, device=device)
            BandCholesky._buffers[wkey] = ws_solve
        self.ws_solve = ws_solve
        self.token = None
        self.ldlt = False

    def _scatter(self, A):
        nat.call("fx_band_from_csr_f64", nat.stream(), self.n, nat.ptr(A.indptr), nat.ptr(A.indices),
                 nat.ptr(A.data), nat.ptr(self.inv), self.bw, self.ldab, nat.ptr(self.ab), nat.ptr(self.ws),
                 self.ws.numel())

    def factor(self, A, token=None):
        """Factor the symmetric FP64 CSRMatrix A: Cholesky, or L D L^T
        (no pivoting) when a pivot is not positive (an indefinite tangent)."""
        import torch
        if A.data.dtype != torch.float64 or A.form != "csr":
            raise NotImplementedError("the direct solver factors an FP64 CSR matrix")
        self.ab.owner = None
        self.token = None
        info = ctypes.c_int32(0)
        self._scatter(A)
        nat.call("fx_band_cholesky_f64", nat.stream(), self.n, self.bw, self.ldab, nat.ptr(self.ab),
                 ctypes.byref(info), nat.ptr(self.ws), self.ws.numel())
        self.ldlt = info.value != 0
        if self.ldlt:
            self._scatter(A)
            nat.call("fx_band_ldlt_f64", nat.stream(), self.n, self.bw, self.ldab, nat.ptr(self.ab),
                     ctypes.byref(info), nat.ptr(self.ws), self.ws.numel())
            if info.value:
                raise RuntimeError(f"direct solver: zero pivot {info.value} of {self.n} in the L D L^T factorisation "
                                   "(J_bc is singular)")
        nat.call("fx_band_invert_diag_f64", nat.stream(), self.n, self.bw, self.ldab,
                 nat.FX_BAND_LDLT if self.ldlt else 0, nat.ptr(self.ab), nat.ptr(self.inv_blocks))
        self.token = token if token is not None else object()
        self.ab.owner = self.token
        self.inv_blocks.owner = self.token
        return self

    @classmethod
    def release(cls):
        """Drop the cached band storage (held for reuse until the process ends)."""
        cls._buffers.clear()

    def valid(self):
        return (self.token is not None and getattr(self.ab, "owner", None) is self.token
                and getattr(self.inv_blocks, "owner", None) is self.token)

    def solve(self, b, out=None, blocked_inverse=True):
        """x = J^-1 b from the factor.  blocked_inverse: the triangular solves
        multiply by the inverses of the diagonal blocks (two launches per
        block, fx_band_solve_inv_f64); False: the sub-block solves of
        fx_band_solve_f64 (same result to rounding)."""
        import torch
        if not self.valid():
            raise RuntimeError("direct solver: the factorisation was overwritten by a later one")
        b = b.to(torch.float64).contiguous()
        x = torch.empty_like(b) if out is None else out
        flags = nat.FX_BAND_LDLT if self.ldlt else 0
        if blocked_inverse:
            nat.call("fx_band_solve_inv_f64", nat.stream(), self.n, self.bw, self.ldab, flags, nat.ptr(self.ab),
                     nat.ptr(self.inv_blocks), nat.ptr(self.perm), nat.ptr(b), nat.ptr(x), nat.ptr(self.ws_solve),
                     self.ws_solve.numel())
        else:
            nat.call("fx_band_solve_f64", nat.stream(), self.n, self.bw, self.ldab, flags, nat.ptr(self.ab),
                     nat.ptr(self.perm), nat.ptr(b), nat.ptr(x), nat.ptr(self.ws), self.ws.numel())
        return x


def solve_direct(problem, ts, tp, bc, u0, opts, symmetric=True, stats=None, keep=None):
    """u = u0 + J_bc^-1 (-R_bc(u0)) by the banded Cholesky (no initial
    iterate: DirectSolverOptions.uses_x0() is False)."""
    import torch
    from ..assembler import assemble_csr, residual
    from .common import check_convergence
    if not symmetric:
        raise NotImplementedError("the direct solver factors the symmetric elimination's SPD J_bc "
                                  "(symmetric_elimination=True)")
    J = assemble_csr(problem, ts, tp, u0, bc, True, torch.float64)
    b = residual(problem, ts, tp, u0, bc, negate=True)
    F = BandCholesky(problem, ts.device)
    token = object()
    F.factor(J, token)
    du = F.solve(b)
    if keep is not None:
        keep.update(form="direct", A=J, dinv=None, symmetric=True, factor=F)
    if opts.check_convergence:
        du = check_convergence(J, du, b, opts)
    if stats is not None:
        stats.iters, stats.rs, stats.outer = 0, 0.0, 1
        stats.inner_iters = []
    u = torch.empty_like(u0)
    nat.call("fx_axpby_f64", nat.stream(), u.numel(), 1.0, nat.ptr(u0), 1.0, nat.ptr(du), nat.ptr(u))
    return u


def direct_adjoint(problem, ts, keep, rhs):
    """lambda = J_bc^-1 rhs with the forward's factor when it is still held
    (J_bc is symmetric), else a fresh factorisation of the kept J_bc."""
    F = keep.get("factor")
    if F is None or not F.valid():
        F = BandCholesky(problem, ts.device).factor(keep["A"])
        keep["factor"] = F
    return F.solve(rhs)
