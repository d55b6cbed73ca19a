"""Krylov building blocks on the device (feax/solvers/common.py API).

create_iterative_solve_fn(options) -> solve(A, b, x0) runs the HIP
Jacobi-PCG (fx_pcg_f64 / fx_pcg_f32) with jax.scipy.sparse.linalg.cg
stopping semantics (feax/solvers/common.py:213-239), or BiCGSTAB
(fx_bicgstab_f64, jax bicgstab semantics, common.py:241-252) for
solver="bicgstab", or restarted GMRES (fx_gmres_f64, jax gmres semantics,
common.py:254-265) for solver="gmres".  Unlike feax's
matrix-free default, ``use_jacobi_preconditioner`` is honoured (the
operator here is the assembled CSR).
"""

import ctypes

import numpy as np

from .. import _native as nat
from ..csr import CSRMatrix

_WS = {}


def _workspace(n, dtype, device):
    """Cached device scratch for the PCG driver (sized by fx_pcg_workspace_bytes)."""
    import torch
    esz = 8 if dtype == torch.float64 else 4
    need = int(nat.lib().fx_pcg_workspace_bytes(n, esz))
    key = (str(device), esz)
    ws = _WS.get(key)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


class PCGInfo:
    def __init__(self, iters, rs):
        self.iters = int(iters)
        self.rs = float(rs)

    def __repr__(self):
        return f"PCGInfo(iters={self.iters}, rs={self.rs:.3e})"


class SpmvTimer:
    """Collects the SpMV kernel times of every PCG solve run while active
    (FX_PCG_TIME: HIP events around each SpMV launch on the solver's
    stream).  ``kernels`` maps a kernel label ("f64/block3", "f32f64/diff3",
    ...) to [total microseconds, launches]."""

    _active = None

    def __init__(self):
        self.kernels = {}

    def __enter__(self):
        SpmvTimer._active = self
        return self

    def __exit__(self, *exc):
        SpmvTimer._active = None
        return False

    def add(self, label, mean_us, count):
        tot = self.kernels.setdefault(label, [0.0, 0])
        tot[0] += mean_us * count
        tot[1] += int(count)

    def mean_us(self, label):
        tot, cnt = self.kernels.get(label, (0.0, 0))
        return tot / cnt if cnt else None


def pcg(A: CSRMatrix, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, dinv=None, block3=None,
        check_every=32, out=None, vec_dtype=None, record=None, use_map=None):
    """Jacobi-PCG on the device; returns (x, PCGInfo).

    Vectors are in the matrix precision, except for a difference-form FP32
    matrix (A.form == "diff3") or vec_dtype=float64 with an FP32 matrix,
    which run FP64 vectors (fx_pcg_f32f64).  record: a KrylovRecord that
    receives the search directions (for recycle()).  use_map: run the
    block-3 SpMV through the matrix's block-column map (fx_pcg_b3map_*):
    None = where it measured faster (FP32 values, FP64 vectors), True =
    whenever the matrix has one, False = never (same results either way).
    """
    import torch
    n = A.shape[0]
    mdt = A.data.dtype
    if A.form in ("diff3", "sym3"):
        vec_dtype = torch.float64
    vdt = mdt if vec_dtype is None else vec_dtype
    b = b.to(vdt).contiguous()
    x = torch.empty(n, dtype=vdt, device=A.data.device) if out is None else out
    x0 = None if x0 is None else x0.to(vdt).contiguous()
    dinv = None if dinv is None else dinv.to(vdt).contiguous()
    if A.form in ("diff3", "sym3"):
        flags = nat.FX_PCG_DIFF3
    else:
        use_b3 = A.block3 if block3 is None else (bool(block3) and A.block3)
        flags = nat.FX_PCG_BLOCK3 if use_b3 else 0
    bcol = A.bcol if flags else None
    timer = SpmvTimer._active
    if timer is not None:
        flags |= nat.FX_PCG_TIME
    ws = _workspace(n, vdt, A.data.device)
    iters = ctypes.c_int32(0)
    rs = (ctypes.c_double * 3)()
    if mdt == torch.float64:
        name = "fx_pcg_f64"
    elif vdt == torch.float64:
        name = "fx_pcg_f32f64"
    else:
        name = "fx_pcg_f32"
    rec = (nat.ptr(record.P) if record is not None else None, nat.ptr(record.pq) if record is not None else None,
           record.capacity if record is not None else 0)
    # A bitwise-symmetric difference form runs its SpMV on upper-block
    # storage (half the matrix bytes; lower blocks are read transposed from
    # blocks the same XCD fetched just before): 78 -> 50 us at 1M DOF.
    upper = A.upper_storage() if (A.form == "sym3" or (use_map is None and (
        (name == "fx_pcg_f32f64" and flags & nat.FX_PCG_DIFF3) or (name == "fx_pcg_f64" and flags & nat.FX_PCG_BLOCK3)))
    ) else None
    if upper is not None:
        sym, vals = upper
        nat.call("fx_pcg_b3sym_f64" if name == "fx_pcg_f64" else "fx_pcg_b3sym_f32f64", nat.stream(), n, A.nse, flags, nat.ptr(A.indptr), ctypes.addressof(sym.struct),
                 nat.ptr(vals), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0), float(tol), float(atol), int(maxiter),
                 int(check_every), nat.ptr(x), ctypes.addressof(iters), ctypes.addressof(rs), *rec, nat.ptr(ws),
                 ws.numel())
        if record is not None:
            record.count = min(iters.value, record.capacity)
        if timer is not None:
            timer.add("f64/sym3" if name == "fx_pcg_f64" else "f32f64/sym3", rs[1], rs[2])
        return x, PCGInfo(iters.value, rs[0])
    # FP32 block-3 values with FP64 vectors: the SpMV with the block-column
    # map (x staged in LDS), same results as the bcol path and 7 % faster at
    # 1M DOF; with FP64 values (32 lanes per node, 8 nodes per workgroup)
    # the staging measured slower, so those keep the bcol path
    want = (name == "fx_pcg_f32f64") if use_map is None else (bool(use_map) and name != "fx_pcg_f32")
    # 256-thread workgroups: 512 (half the SpMV partials to sum) measured slower
    bmap = A.b3map(A.data.element_size(), 256) if (
        want and flags & (nat.FX_PCG_BLOCK3 | nat.FX_PCG_DIFF3) and A.b3map is not None) else None
    if bmap is not None:
        nat.call(name.replace("fx_pcg_", "fx_pcg_b3map_"), nat.stream(), n, A.nse, flags, nat.ptr(A.indptr),
                 ctypes.addressof(bmap.struct), nat.ptr(A.data), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0), float(tol),
                 float(atol), int(maxiter), int(check_every), nat.ptr(x), ctypes.addressof(iters),
                 ctypes.addressof(rs), *rec, nat.ptr(ws), ws.numel())
    else:
        nat.call(name, nat.stream(), n, A.nse, flags, nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(bcol),
                 nat.ptr(A.data), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0), float(tol), float(atol), int(maxiter),
                 int(check_every), nat.ptr(x), ctypes.addressof(iters), ctypes.addressof(rs), *rec, nat.ptr(ws),
                 ws.numel())
    if record is not None:
        record.count = min(iters.value, record.capacity)
    if timer is not None:
        form = "diff3" if flags & nat.FX_PCG_DIFF3 else ("block3" if flags & nat.FX_PCG_BLOCK3 else "csr")
        timer.add(f"{name[len('fx_pcg_'):]}/{form}", rs[1], rs[2])
    return x, PCGInfo(iters.value, rs[0])


_BCG_WS = {}


def bicgstab(A: CSRMatrix, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, dinv=None, block3=None,
             check_every=16, out=None):
    """BiCGSTAB on the device (fx_bicgstab_f64, FP64) with the iteration and
    stopping rule of jax.scipy.sparse.linalg.bicgstab (feax/solvers/
    common.py:241-252); M = Jacobi (dinv) or none.  Returns (x, PCGInfo)."""
    import torch
    if A.form != "csr" or A.data.dtype != torch.float64:
        raise NotImplementedError("bicgstab runs on an FP64 CSR operator")
    n = A.shape[0]
    b = b.to(torch.float64).contiguous()
    x = torch.empty(n, dtype=torch.float64, device=A.data.device) if out is None else out
    x0 = None if x0 is None else x0.to(torch.float64).contiguous()
    dinv = None if dinv is None else dinv.to(torch.float64).contiguous()
    use_b3 = A.block3 if block3 is None else (bool(block3) and A.block3)
    flags = nat.FX_PCG_BLOCK3 if use_b3 else 0
    need = int(nat.lib().fx_bicgstab_workspace_bytes(n))
    key = str(A.data.device)
    ws = _BCG_WS.get(key)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=A.data.device)
        _BCG_WS[key] = ws
    iters = ctypes.c_int32(0)
    rs = ctypes.c_double(0.0)
    nat.call("fx_bicgstab_f64", nat.stream(), n, A.nse, flags, nat.ptr(A.indptr), nat.ptr(A.indices),
             nat.ptr(A.bcol if use_b3 else None), nat.ptr(A.data), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0),
             float(tol), float(atol), int(maxiter), int(check_every), nat.ptr(x), ctypes.byref(iters),
             ctypes.byref(rs), nat.ptr(ws), ws.numel())
    return x, PCGInfo(iters.value, rs.value)


_GM_WS = {}


def gmres(A: CSRMatrix, b, x0=None, tol=1e-10, atol=1e-10, restart=20, maxiter=10000, dinv=None, block3=None,
          check_every=1, out=None):
    """Restarted GMRES on the device (fx_gmres_f64, FP64) with the iteration
    and stopping rule of jax.scipy.sparse.linalg.gmres, solve_method=
    'batched' (feax/solvers/common.py:254-265): left preconditioning with
    M = Jacobi (dinv) or none, maxiter counts restart cycles.  Returns
    (x, PCGInfo) with iters = cycles and rs = ||M (b - A x)||^2."""
    import torch
    if A.form != "csr" or A.data.dtype != torch.float64:
        raise NotImplementedError("gmres runs on an FP64 CSR operator")
    restart = 20 if restart is None else int(restart)
    if not 1 <= restart <= 64:
        raise ValueError(f"gmres restart must lie in [1, 64], got {restart}")
    n = A.shape[0]
    b = b.to(torch.float64).contiguous()
    x = torch.empty(n, dtype=torch.float64, device=A.data.device) if out is None else out
    x0 = None if x0 is None else x0.to(torch.float64).contiguous()
    dinv = None if dinv is None else dinv.to(torch.float64).contiguous()
    use_b3 = A.block3 if block3 is None else (bool(block3) and A.block3)
    flags = nat.FX_PCG_BLOCK3 if use_b3 else 0
    need = int(nat.lib().fx_gmres_workspace_bytes(n, restart, flags))
    key = str(A.data.device)
    ws = _GM_WS.get(key)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=A.data.device)
        _GM_WS[key] = ws
    cycles = ctypes.c_int32(0)
    rn = ctypes.c_double(0.0)
    nat.call("fx_gmres_f64", nat.stream(), n, A.nse, flags, nat.ptr(A.indptr), nat.ptr(A.indices),
             nat.ptr(A.bcol if use_b3 else None), nat.ptr(A.data), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0),
             float(tol), float(atol), restart, int(maxiter), int(check_every), nat.ptr(x), ctypes.byref(cycles),
             ctypes.byref(rn), nat.ptr(ws), ws.numel())
    return x, PCGInfo(cycles.value, rn.value * rn.value)


def krylov(A, b, x0, opts, dinv=None, block3=None, out=None):
    """The Krylov method opts.solver names ("cg", "bicgstab" or "gmres") on A."""
    if opts.solver == "gmres":
        return gmres(A, b, x0, opts.tol, opts.atol, opts.restart, opts.maxiter, dinv, block3, out=out)
    if opts.solver == "bicgstab":
        return bicgstab(A, b, x0, opts.tol, opts.atol, opts.maxiter, dinv, block3, min(opts.check_every, 16),
                        out=out)
    return pcg(A, b, x0, opts.tol, opts.atol, opts.maxiter, dinv, block3, opts.check_every, out=out)


def pcg_matfree(problem, ts, tp, u_lin, bc, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, check_every=32,
                out=None, dinv=None):
    """CG on feax's matrix-free operator J_bc(u_lin) (fx_pcg_matfree_f64);
    returns (x, PCGInfo).  u_lin: linearisation point (Neo-Hookean; None for
    linear laws)."""
    import torch
    from ..assembler import _as_dev, resolve_material
    mat, p = resolve_material(problem, tp, ts)
    n = ts.n
    b = b.to(torch.float64).contiguous()
    x = torch.empty(n, dtype=torch.float64, device=ts.device) if out is None else out
    x0 = None if x0 is None else x0.to(torch.float64).contiguous()
    u = _as_dev(u_lin, ts) if u_lin is not None else None
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    flags = nat.FX_PCG_TIME if SpmvTimer._active is not None else 0
    ws = _workspace(n, torch.float64, ts.device)
    iters = ctypes.c_int32(0)
    rs = (ctypes.c_double * 3)()
    nat.call("fx_pcg_matfree_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat), nat.ptr(p),
             nat.ptr(u), nat.ptr(mask), flags, nat.ptr(dinv), nat.ptr(b), nat.ptr(x0), float(tol), float(atol),
             int(maxiter), int(check_every), nat.ptr(x), ctypes.addressof(iters), ctypes.addressof(rs), nat.ptr(ws),
             ws.numel())
    if flags:
        SpmvTimer._active.add("f64/matfree", rs[1], rs[2])
    return x, PCGInfo(iters.value, rs[0])


class KrylovRecord:
    """Device storage for the search directions of one CG solve (FP32,
    capacity x n) and their curvatures p.Ap, used to start a later solve
    with the same matrix from the Galerkin projection of its right-hand
    side (fx_krylov_project_f32)."""

    def __init__(self, n, capacity, device):
        import torch
        self.n, self.capacity, self.count = int(n), int(capacity), 0
        self.P = torch.empty((self.capacity, self.n), dtype=torch.float32, device=device)
        self.pq = torch.empty(self.capacity, dtype=torch.float64, device=device)
        self._ws = torch.empty(int(nat.lib().fx_krylov_project_workspace_bytes(self.capacity)), dtype=torch.uint8,
                               device=device)

    def project(self, r, out=None):
        """x0 = sum_k p_k (p_k . r) / (p_k . A p_k) over the recorded directions."""
        import torch
        r = r.to(torch.float64).contiguous()
        x0 = torch.empty(self.n, dtype=torch.float64, device=r.device) if out is None else out
        nat.call("fx_krylov_project_f32", nat.stream(), self.n, self.count, nat.ptr(self.P), nat.ptr(self.pq),
                 nat.ptr(r), nat.ptr(x0), nat.ptr(self._ws), self._ws.numel())
        return x0


def create_jacobi_preconditioner(A: CSRMatrix, shift=1e-12):
    """dinv = 1 / (diag + shift) (feax/solvers/common.py:86-94), as a device vector."""
    import torch
    d = torch.empty(A.shape[0], dtype=A.data.dtype, device=A.data.device)
    name = "fx_csr_jacobi_f64" if A.data.dtype == torch.float64 else "fx_csr_jacobi_f32"
    nat.call(name, nat.stream(), A.shape[0], nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(A.data),
             float(shift), nat.ptr(d))
    return d


def inner_tolerance(opts, target, rn):
    """Relative tolerance of a mixed-precision inner solve whose right-hand
    side has norm rn when the outer loop stops at target: inner_tol, or
    the smaller reduction the target still needs (times inner_safety)."""
    if opts.inner_safety <= 0 or rn <= 0:
        return opts.inner_tol
    return min(0.5, max(opts.inner_tol, opts.inner_safety * target / rn))


def create_x0(bc_rows=None, bc_vals=None):
    """x0_fn(current_sol) lifting BC increments (feax/solvers/common.py:23-53)."""
    import torch

    def x0_fn(current_sol):
        x0 = torch.empty_like(current_sol, dtype=torch.float64)
        if bc_rows is None or bc_vals is None or len(bc_rows) == 0:
            nat.call("fx_bc_lift_f64", nat.stream(), x0.numel(), 0, None, None, None, nat.ptr(x0))
            return x0
        rows = torch.as_tensor(np.asarray(bc_rows, np.int32), device=x0.device) \
            if not isinstance(bc_rows, torch.Tensor) else bc_rows.to(torch.int32)
        vals = torch.as_tensor(np.asarray(bc_vals, np.float64), device=x0.device) \
            if not isinstance(bc_vals, torch.Tensor) else bc_vals.detach().to(device=x0.device, dtype=torch.float64)
        u0 = current_sol.to(torch.float64).contiguous()
        nat.call("fx_bc_lift_f64", nat.stream(), x0.numel(), len(rows), nat.ptr(rows.contiguous()),
                 nat.ptr(vals.contiguous()), nat.ptr(u0), nat.ptr(x0))
        return x0
    return x0_fn


def device_norm(v):
    """||v||_2 in FP64 with the fixed-order device reduction."""
    import torch
    v = v.to(torch.float64).contiguous()
    out = torch.empty(1, dtype=torch.float64, device=v.device)
    ws = _dot_ws(v.device)
    nat.call("fx_dot_f64", nat.stream(), v.numel(), nat.ptr(v), nat.ptr(v), nat.ptr(out), nat.ptr(ws), ws.numel())
    return float(out.cpu()[0]) ** 0.5


_DOT_WS = {}


def _dot_ws(device):
    import torch
    k = str(device)
    if k not in _DOT_WS:
        _DOT_WS[k] = torch.empty(8 * 2048 + 256, dtype=torch.uint8, device=device)
    return _DOT_WS[k]


def check_convergence(A, x, b, solver_options, matrix_view=None, solver_label="Linear solver"):
    """NaN-fill unless ||Ax - b|| / (||b|| + 1e-12) < threshold
    (feax/solvers/common.py:376-395).  With an UPPER or LOWER matrix_view,
    A x is the product with the full symmetric matrix A + A^T - diag(A)
    (feax/solvers/common.py:364-373)."""
    import torch
    from ..csr import expand_view
    if isinstance(A, CSRMatrix):
        Ax = expand_view(A, matrix_view) @ x
    else:
        Ax = A(x) if callable(A) else A @ x
    r = torch.empty_like(b, dtype=torch.float64)
    nat.call("fx_axpby_f64", nat.stream(), r.numel(), 1.0, nat.ptr(Ax.to(torch.float64).contiguous()), -1.0,
             nat.ptr(b.to(torch.float64).contiguous()), nat.ptr(r))
    rn = device_norm(r)
    rel = rn / (device_norm(b) + 1e-12)
    if getattr(solver_options, "verbose", False):
        print(f"{solver_label} residual: abs={rn:.6e}, rel={rel:.6e}")
    if rel < getattr(solver_options, "convergence_threshold", 0.1):
        return x
    return torch.full_like(x, float("nan"))


def create_iterative_solve_fn(options):
    """solve(A, b, x0) -> x for KrylovSolverOptions (common.py:213-270)."""
    if options.solver == "auto":
        raise ValueError("KrylovSolverOptions.solver is 'auto' (unresolved). "
                         "Call resolve_iterative_solver() before create_iterative_solve_fn().")
    if options.solver not in ("cg", "bicgstab", "gmres"):
        raise NotImplementedError(f"{options.solver} is not on the HIP path")
    if options.preconditioner is not None:
        raise NotImplementedError("custom preconditioners are not on the HIP path; GeometricMultigrid needs the "
                                  "problem and runs through create_solver / create_linear_solver")

    def solve(A, b, x0):
        dinv = create_jacobi_preconditioner(A, options.jacobi_shift) if options.use_jacobi_preconditioner else None
        b3 = {"auto": None, "csr": False, "block3": True}[options.spmv]
        x, info = krylov(A, b, x0, options, dinv, b3)
        solve.last_info = info
        return x
    solve.last_info = None
    return solve


class _Pattern:
    """A CSR pattern standing in for a Problem in band_ordering (one DOF per
    graph node)."""

    def __init__(self, A):
        self.csr_indptr = A.indptr.detach().cpu().numpy()
        self.csr_indices = A.indices.detach().cpu().numpy()
        self.vec = 1


def create_direct_solve_fn(options, *, cache_namespace="global"):
    """solve(A, b, x0=None) -> x for DirectSolverOptions
    (feax/solvers/common.py:115-210): the device banded factorisation of the
    symmetric FP64 CSRMatrix A (Cholesky, or L D L^T without pivoting),
    ordered by reverse Cuthill-McKee on A's own graph; the factorisation of
    the last pattern is reused while A's values are unchanged only under
    reuse_factorization."""
    from .direct import BandCholesky
    if options.solver == "auto":
        raise ValueError("DirectSolverOptions.solver is 'auto' (unresolved). "
                         "Call resolve_direct_solver() before create_direct_solve_fn().")
    state = {}

    def solve(A, b, x0=None):
        key = (A.indptr.data_ptr(), A.indices.data_ptr(), A.shape)
        if state.get("key") != key:
            state.update(key=key, chol=BandCholesky(_Pattern(A), A.data.device), data=None)
        chol = state["chol"]
        same = options.reuse_factorization and state["data"] is A.data and chol.valid()
        if not same:
            chol.factor(A)
            state["data"] = A.data
        return chol.solve(b)
    return solve


def create_linear_solve_fn(solver_options, *, cache_namespace="global"):
    """The solve(A, b, x0) of DirectSolverOptions or KrylovSolverOptions
    (feax/solvers/common.py:273-296)."""
    from .options import DirectSolverOptions, KrylovSolverOptions
    if isinstance(solver_options, DirectSolverOptions):
        return create_direct_solve_fn(solver_options, cache_namespace=cache_namespace)
    if isinstance(solver_options, KrylovSolverOptions):
        return create_iterative_solve_fn(solver_options)
    raise TypeError("Unsupported solver option type. Expected DirectSolverOptions or KrylovSolverOptions, got "
                    f"{type(solver_options).__name__}.")
