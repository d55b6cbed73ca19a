"""Batched uneven instances: many HEX8 problems solved in one set of launches.

feax batches solves with jax.vmap over the traced parameters of one mesh
(tests/test_vmap.py:63-106, examples/benchmark/bench_linear_elasticity.py:
416-459); it cannot batch different meshes.  Here instances -- meshes of
any size, each with its own parameters, loads and Dirichlet values -- are
concatenated into one block-diagonal problem:

  BatchedProblem(problems)          nodes, cells and loaded faces stacked,
                                    feax's CSR pattern of the union (the
                                    instances' patterns side by side)
  batch_bc(bp, bcs)                 Dirichlet rows with node offsets
  batch_params(bp, tps)             volume / surface vars concatenated
  create_batched_solver(bp, bc, options)
                                    solver(tp) -> u (concatenated); one
                                    assembly launch for all instances,
                                    segmented PCG (fx_pcg_batched_*): per
                                    instance CG scalars, stopping tests and
                                    convergence flags, iteration chunks
                                    replayed as a HIP graph, SpMV lane count
                                    autotuned per shape
  bp.split(u)                       per-instance views

The same machinery batches parameter sets of one mesh (pass the same
problem several times), the vmap use of feax.
"""

import numpy as np

from . import _native as nat
from .DCboundary import DirichletBC
from .fe import NUM_FACE_QUADS
from .mesh import Mesh
from .problem import FiniteElement, MatrixView, Problem
from .traced_params import TracedParams


def _same(a, b):
    """Same law / map: same type and equal attributes (arrays compared by value)."""
    if type(a) is not type(b):
        return False
    va, vb = vars(a), vars(b)
    if va.keys() != vb.keys():
        return False
    return all(np.array_equal(np.asarray(va[k]), np.asarray(vb[k])) for k in va)


class BatchedProblem(Problem):
    """Block-diagonal concatenation of independent HEX8 problems that share
    the material law and the list of surface maps."""

    def __init__(self, problems):
        problems = list(problems)
        if not problems:
            raise ValueError("BatchedProblem needs at least one problem")
        p0 = problems[0]
        nb = len(p0.boundary_inds_list)
        tm0 = p0.get_tensor_map() or p0.get_energy_density()
        maps0 = p0.get_surface_maps() or []
        for p in problems[1:]:
            tm = p.get_tensor_map() or p.get_energy_density()
            if not _same(tm, tm0):
                raise ValueError("batched instances must share the material law")
            if len(p.boundary_inds_list) != nb:
                raise ValueError("batched instances must have the same number of loaded boundaries")
            maps = p.get_surface_maps() or []
            if len(maps) != len(maps0) or not all(_same(a, b) for a, b in zip(maps, maps0)):
                raise ValueError("batched instances must share the surface maps")
        self.instances = problems
        self._tensor_map, self._surface_maps = tm0, maps0
        self._energy = p0.get_energy_density() is not None and p0.get_tensor_map() is None
        self.vec, self.dim, self.ele_type, self.gauss_order, self.npe = 3, 3, "HEX8", None, 8
        self.location_fns, self.matrix_view, self.additional_info = None, MatrixView.FULL, ()
        self.nseg = len(problems)
        self.seg_node_ptr = np.cumsum([0] + [p.num_total_nodes for p in problems]).astype(np.int32)
        self.seg_cell_ptr = np.cumsum([0] + [p.num_cells for p in problems]).astype(np.int64)
        points = np.concatenate([p.points for p in problems])
        cells = np.concatenate([p.cells.astype(np.int64) + off for p, off in zip(problems, self.seg_node_ptr[:-1])])
        self.mesh = Mesh(points, cells.astype(np.int32))
        self.num_vars = 1
        self.fes = [FiniteElement(self.mesh, 3, 3)]
        self.points = np.ascontiguousarray(points)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.cells_list = [self.cells]
        self.num_cells = len(self.cells)
        self.num_total_nodes = len(self.points)
        self.num_total_dofs_all_vars = 3 * self.num_total_nodes
        self.offset = [0]
        # boundary j = the instances' boundary j faces, instance by instance
        self.boundary_inds_list = []
        self.seg_face_ptr = []
        for j in range(nb):
            parts, counts = [], [0]
            for p, coff in zip(problems, self.seg_cell_ptr[:-1]):
                b = np.asarray(p.boundary_inds_list[j], np.int64).reshape(-1, 2).copy()
                b[:, 0] += coff
                parts.append(b)
                counts.append(len(b))
            self.boundary_inds_list.append(np.concatenate(parts) if parts else np.zeros((0, 2), np.int64))
            self.seg_face_ptr.append(np.cumsum(counts))
        self._build_csr_assembly_structure()

    def get_tensor_map(self):
        return None if self._energy else self._tensor_map

    def get_energy_density(self):
        return self._tensor_map if self._energy else None

    def get_surface_maps(self):
        return self._surface_maps

    def get_mass_map(self):
        return None

    def split(self, u):
        """Per-instance views of a concatenated DOF vector."""
        u = getattr(u, "dofs", u)
        return [u[3 * a:3 * b] for a, b in zip(self.seg_node_ptr[:-1], self.seg_node_ptr[1:])]


def batch_bc(bp, bcs):
    """One DirichletBC for the batch: the instances' rows shifted by their
    DOF offsets (sorted, since instances are stacked in order)."""
    rows, vals, mask = [], [], []
    for bc, off in zip(bcs, bp.seg_node_ptr[:-1]):
        rows.append(np.asarray(bc.bc_rows, np.int64) + 3 * int(off))
        vals.append(bc.vals_numpy())
        mask.append(np.asarray(bc.bc_mask, bool))
    return DirichletBC(np.concatenate(rows).astype(np.int32), np.concatenate(mask), np.concatenate(vals),
                       bp.num_total_dofs_all_vars)


def batch_params(bp, tps):
    """One TracedParams for the batch: volume_vars[0] concatenated (node or
    cell variables), surface vars concatenated per boundary (each (F_i, 4)
    or (F_i,) per instance)."""
    import torch
    vol = []
    for tp, p in zip(tps, bp.instances):
        v = tp.volume_vars[0]
        v = v.detach().to(torch.float64).reshape(-1) if isinstance(v, torch.Tensor) else \
            torch.as_tensor(np.asarray(v, np.float64).reshape(-1))
        if v.numel() == 1:
            v = v.expand(p.num_cells)
        vol.append(v.to("cpu"))
    surf = []
    for j in range(len(bp.boundary_inds_list)):
        parts = []
        for tp, p in zip(tps, bp.instances):
            nf = len(p.boundary_inds_list[j])
            sv = tp.surface_vars[j][0] if (tp.surface_vars and len(tp.surface_vars) > j) else np.zeros((nf, 4))
            sv = sv.detach().cpu().numpy() if isinstance(sv, torch.Tensor) else np.asarray(sv, np.float64)
            if nf == 0:
                continue
            if sv.size == 1:
                sv = np.full(nf, float(sv.reshape(())))
            parts.append(np.broadcast_to(sv.reshape(nf, -1), (nf, NUM_FACE_QUADS)))
        surf.append((np.concatenate(parts) if parts else np.zeros((0, NUM_FACE_QUADS)),))
    return TracedParams(volume_vars=(torch.cat(vol).numpy(),), surface_vars=surf)


class BatchedStats:
    def __init__(self):
        self.iters = []            # per instance, summed over outer steps
        self.outer = 0
        self.relres = []           # per instance, FP64 ||b - K u|| / ||b||
        self.lanes = None          # SpMV lanes per node of the last solve
        self.spmv = None           # its SpMV configuration: lanes, threads, map

    def __repr__(self):
        return (f"BatchedStats(outer={self.outer}, spmv={self.spmv}, iters=[{min(self.iters)}..{max(self.iters)}], "
                f"relres max {max(self.relres):.2e})")


def _seg_norms(bp, ts, v):
    import torch
    key = "seg_row_ptr"
    if key not in ts._cache:
        ts._cache[key] = torch.as_tensor((3 * bp.seg_node_ptr).astype(np.int32), device=ts.device)
    out = torch.empty(bp.nseg, dtype=torch.float64, device=ts.device)
    nat.call("fx_seg_dot_f64", nat.stream(), bp.nseg, nat.ptr(ts._cache[key]), nat.ptr(v), nat.ptr(v), nat.ptr(out))
    return out.sqrt().cpu().numpy()


def _row_seg(bp, ts):
    import torch
    key = "row_seg"
    if key not in ts._cache:
        counts = 3 * np.diff(bp.seg_node_ptr)
        ts._cache[key] = torch.as_tensor(np.repeat(np.arange(bp.nseg), counts), device=ts.device)
    return ts._cache[key]


_WS = {}


class SegKrylovRecord:
    """Per-instance search directions of one batched inner solve (FP32,
    capacity x n) and their p.A p (capacity x nseg); project(r) is the
    per-instance Galerkin projection of r onto them (fx_seg_project_f32)."""

    def __init__(self, bp, n, capacity, device):
        import torch
        self.bp, self.n, self.capacity = bp, n, int(capacity)
        self.P = torch.empty((self.capacity, n), dtype=torch.float32, device=device)
        self.pq = torch.empty((self.capacity, bp.nseg), dtype=torch.float64, device=device)
        self.counts = np.zeros(bp.nseg, np.int32)

    def project(self, r):
        import torch
        snp = np.ascontiguousarray(self.bp.seg_node_ptr, dtype=np.int32)
        kmax = int(self.counts.max()) if len(self.counts) else 0
        need = int(nat.lib().fx_seg_project_workspace_bytes(self.bp.nseg, nat.ptr(snp), kmax))
        ws = torch.empty(need, dtype=torch.uint8, device=r.device)
        x0 = torch.empty(self.n, dtype=torch.float64, device=r.device)
        nat.call("fx_seg_project_f32", nat.stream(), self.bp.nseg, nat.ptr(snp), kmax, nat.ptr(self.counts),
                 nat.ptr(self.P), nat.ptr(self.pq), nat.ptr(r.to(torch.float64).contiguous()), nat.ptr(x0),
                 nat.ptr(ws), need)
        return x0


_SEG_RECORDS = {}


def _seg_record_for(bp, n, capacity, device):
    key = (str(device), n, bp.nseg)
    rec = _SEG_RECORDS.get(key)
    if rec is None or rec.capacity < capacity or rec.bp is not bp:
        _SEG_RECORDS.pop(key, None)
        rec = SegKrylovRecord(bp, n, capacity, device)
        _SEG_RECORDS[key] = rec
    return rec


def pcg_batched(bp, A, b, x0, dinv, tol, atol, maxiter, check_every=32, out=None, autotune=True, graph=True,
                record=None, spmv=None):
    """Segmented Jacobi-PCG over the instances (fx_pcg_batched_f64 /
    _f32f64; with record, an FP32-matrix solve that keeps each instance's
    search directions in it, fx_pcg_batched_record_f32f64).  Returns (x,
    iters[nseg], rs[nseg], spmv) with spmv the SpMV configuration that ran:
    {"lanes", "threads" (per workgroup), "map" (x staged in LDS through the
    block-column map)}.  spmv given: run that configuration (FX_PCG_SEG_FORCE)
    instead of the default or the autotuner's."""
    import torch
    n = A.shape[0]
    dev = A.data.device
    b = b.to(torch.float64).contiguous()
    x = torch.empty(n, dtype=torch.float64, device=dev) if out is None else out
    x0 = None if x0 is None else x0.to(torch.float64).contiguous()
    dinv = None if dinv is None else dinv.to(torch.float64).contiguous()
    snp = np.ascontiguousarray(bp.seg_node_ptr, dtype=np.int32)
    need = int(nat.lib().fx_pcg_batched_workspace_bytes(bp.nseg, nat.ptr(snp), int(A.nse), 8))
    ws = _WS.get(str(dev))
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=dev)
        _WS[str(dev)] = ws
    flags = (nat.FX_PCG_AUTOTUNE if autotune else 0) | (0 if graph else nat.FX_PCG_NO_GRAPH)
    tuned = np.zeros(3, np.int32)
    if spmv is not None:
        flags |= nat.FX_PCG_SEG_FORCE
        tuned[:] = (int(spmv["lanes"]), int(spmv["threads"]), int(bool(spmv["map"])))
    if A.data.dtype == torch.float64:
        name = "fx_pcg_batched_f64"
    else:
        name = "fx_pcg_batched_f32f64"
        if A.form == "diff3":
            flags |= nat.FX_PCG_DIFF3
    iters = np.zeros(bp.nseg, np.int32)
    rs = np.zeros(bp.nseg, np.float64)
    common = (nat.stream(), bp.nseg, nat.ptr(snp), A.nse, flags, nat.ptr(A.indptr), nat.ptr(A.bcol),
              nat.ptr(A.data), nat.ptr(dinv), nat.ptr(b), nat.ptr(x0), float(tol), float(atol), int(maxiter),
              int(check_every), nat.ptr(x), nat.ptr(iters), nat.ptr(rs), nat.ptr(tuned))
    if record is not None:
        if A.data.dtype != torch.float32:
            raise ValueError("recording is for the FP32-matrix batched solve")
        nat.call("fx_pcg_batched_record_f32f64", *common, nat.ptr(record.P), nat.ptr(record.pq), record.capacity,
                 nat.ptr(ws), ws.numel())
        record.counts = np.minimum(iters, record.capacity).astype(np.int32)
    else:
        nat.call(name, *common, nat.ptr(ws), ws.numel())
    return x, iters, rs, {"lanes": int(tuned[0]), "threads": int(tuned[1]), "map": bool(tuned[2])}


def create_batched_solver(bp, bc, solver_options=None, traced_structure=None, autotune=True, graph=True,
                          spmv=None):
    """solver(traced_params, initial_guess=None) -> concatenated u (FP64).

    precision "fp64": FP64 assembly of all instances in one launch, one
    segmented Jacobi-PCG.  precision "mixed": FP32 difference-form matrix,
    FP64 vectors, and an FP64 defect correction whose per-instance residual
    norms decide which instances take another inner solve.  spmv: a fixed
    SpMV configuration for pcg_batched (default: autotuned when autotune)."""
    import torch
    from .assembler import TracedStructure, assemble_csr, residual
    from .solvers.common import create_x0, inner_tolerance
    from .solvers.options import KrylovSolverOptions
    from .solvers.linear import _dev_vec, _zero_like
    opts = solver_options or KrylovSolverOptions(solver="cg", use_jacobi_preconditioner=True)
    ts = traced_structure or TracedStructure.from_problem(bp)
    stats = BatchedStats()

    def solver(traced_params, initial_guess=None):
        u0 = _dev_vec(initial_guess, ts) if initial_guess is not None else _zero_like(bp, bc, ts)
        shift = opts.jacobi_shift if opts.use_jacobi_preconditioner else None
        x0 = create_x0(bc.bc_rows, bc.bc_vals)(u0)
        if opts.precision == "fp64":
            out = assemble_csr(bp, ts, traced_params, u0, bc, True, torch.float64, jacobi_shift=shift)
            A, dinv = out if shift is not None else (out, None)
            b = residual(bp, ts, traced_params, u0, bc, negate=True)
            du, iters, rs, ran = pcg_batched(bp, A, b, x0, dinv, opts.tol, opts.atol, opts.maxiter,
                                              opts.check_every, autotune=autotune, graph=graph, spmv=spmv)
            u = u0 + du
            stats.iters, stats.outer, stats.spmv = [int(i) for i in iters], 1, ran
            stats.lanes = ran["lanes"]
            bn = _seg_norms(bp, ts, b)
            rn = _seg_norms(bp, ts, residual(bp, ts, traced_params, u, bc, negate=True))
            stats.relres = [float(r / bb) if bb > 0 else 0.0 for r, bb in zip(rn, bn)]
            return u
        out = assemble_csr(bp, ts, traced_params, u0, bc, True, torch.float32, diff_form=True, jacobi_shift=shift)
        A, dinv = out if shift is not None else (out, None)
        u = u0.clone()
        r = residual(bp, ts, traced_params, u, bc, negate=True)
        bn = _seg_norms(bp, ts, r)
        target = np.maximum(opts.tol * bn, opts.atol)
        rn = bn.copy()
        row_seg = _row_seg(bp, ts)
        total = np.zeros(bp.nseg, np.int64)
        d = torch.empty_like(u)
        outer = 0
        # Krylov recycling per instance: the first inner solve records its
        # directions, later ones start from the projection onto them.
        record = _seg_record_for(bp, ts.n, min(opts.recycle, opts.maxiter), ts.device) if opts.recycle > 0 else None
        while np.any(rn > target) and outer < opts.max_outer:
            act = rn > target
            active = torch.as_tensor(act.astype(np.float64), device=ts.device)
            rhs = r * active[row_seg]
            # one tolerance for all instances: the strictest need among the active ones
            tin = min(inner_tolerance(opts, t_, r_) for t_, r_ in zip(target[act], rn[act]))
            x0i = record.project(rhs) if (record is not None and outer > 0) else None
            _, iters, _, ran = pcg_batched(bp, A, rhs, x0i, dinv, tin, 0.0, opts.maxiter, opts.check_every,
                                           out=d, autotune=autotune, graph=graph, spmv=spmv,
                                           record=record if outer == 0 else None)
            u += d
            residual(bp, ts, traced_params, u, bc, negate=True, out=r)
            rn = _seg_norms(bp, ts, r)
            total += iters
            outer += 1
            stats.spmv, stats.lanes = ran, ran["lanes"]
        stats.iters, stats.outer = [int(i) for i in total], outer
        stats.relres = [float(a / bb) if bb > 0 else 0.0 for a, bb in zip(rn, bn)]
        return u

    solver.stats = stats
    solver.traced_structure = ts
    return solver


def create_vmap_solver(problem, bc, solver_options=None, autotune=True, graph=True):
    """solve(traced_params, bc_vals=None) -> (B, n) FP64 tensor: the
    jax.vmap(solver) use of feax (tests/test_vmap.py:63-106,
    tests/test_bc_vmap.py) on one HEX8 mesh.  Every tensor / array of
    traced_params carries a leading batch axis B; bc_vals, when given, is
    (B, len(bc.bc_rows)) (one set of Dirichlet values per lane).  The lanes
    are solved together by the batched engine (one assembly launch, one
    segmented PCG; per-lane stopping tests as in a vmapped jax cg); other
    Krylov methods (bicgstab, gmres) and preconditioned or matrix-free forms
    run one create_solver call per lane."""
    import torch
    from .solvers.options import KrylovSolverOptions
    opts = solver_options or KrylovSolverOptions(solver="cg", use_jacobi_preconditioner=True)
    cache = {}

    def lane(x, i):
        return x[i] if x is not None else None

    lane_solver = []

    def per_lane(traced_params, bc_vals):
        # Krylov methods other than CG (and preconditioned / matrix-free
        # forms): one create_solver call per lane, stacked; autograd runs per
        # lane through the solver's own backward.
        from .solver import create_solver
        if not lane_solver:
            lane_solver.append(create_solver(problem, bc, opts, linear=True, return_solution=False))
        leaves = list(traced_params.volume_vars) + [v for sv in traced_params.surface_vars for v in sv]
        B = int(leaves[0].shape[0]) if leaves else int(bc_vals.shape[0])
        out = []
        for i in range(B):
            tpi = TracedParams(volume_vars=tuple(lane(v, i) for v in traced_params.volume_vars),
                               surface_vars=[tuple(lane(v, i) for v in sv) for sv in traced_params.surface_vars])
            bci = bc if bc_vals is None else bc.replace_vals(bc_vals[i])
            out.append(lane_solver[0](tpi, None, bc=bci))
        return torch.stack(out)

    def run(traced_params, bc_vals, keep=None):
        leaves = list(traced_params.volume_vars) + [v for sv in traced_params.surface_vars for v in sv]
        sizes = {int(v.shape[0]) for v in leaves if hasattr(v, "shape") and len(v.shape)}
        if bc_vals is not None:
            sizes.add(int(bc_vals.shape[0]))
        if len(sizes) != 1:
            raise ValueError(f"every batched input needs the same leading axis (got {sorted(sizes)})")
        B = sizes.pop()
        if B not in cache:
            from .assembler import TracedStructure
            bp = BatchedProblem([problem] * B)
            cache[B] = (bp, TracedStructure.from_problem(bp))
        bp, ts = cache[B]
        vals = bc_vals.detach().cpu().numpy() if isinstance(bc_vals, torch.Tensor) else bc_vals
        bcs = [bc if vals is None else bc.replace_vals(np.asarray(vals[i], np.float64)) for i in range(B)]
        tps = [TracedParams(volume_vars=tuple(lane(v, i) for v in traced_params.volume_vars),
                            surface_vars=[tuple(lane(v, i) for v in sv) for sv in traced_params.surface_vars])
               for i in range(B)]
        solver = create_batched_solver(bp, batch_bc(bp, bcs), opts, traced_structure=ts, autotune=autotune,
                                       graph=graph)
        tpb = batch_params(bp, tps)
        tpd = TracedParams(volume_vars=tuple(torch.as_tensor(v, device=ts.device) for v in tpb.volume_vars),
                           surface_vars=[tuple(torch.as_tensor(v, device=ts.device) for v in sv)
                                         for sv in tpb.surface_vars])
        u = solver(tpd)
        solve.stats = solver.stats
        if keep is not None:
            keep.update(bp=bp, ts=ts, bc=batch_bc(bp, bcs), tp=tpd)
        return u.reshape(B, -1)

    def solve(traced_params, bc_vals=None):
        from .solvers.adjoint import _leaves, _needs_grad
        if opts.solver != "cg" or opts.precision not in ("fp64", "mixed") or opts.operator != "csr" \
                or getattr(opts, "preconditioner", None) is not None:
            return per_lane(traced_params, bc_vals)
        if not _needs_grad(traced_params):
            return run(traced_params, bc_vals)
        leaves = _leaves(traced_params)
        return _vmap_function().apply(dict(run=run, tp=traced_params, bc_vals=bc_vals, opts=opts,
                                           paths=[q for q, _ in leaves]), *[t for _, t in leaves])

    return solve


_VMAP_FN = None


def _vmap_function():
    """autograd of the vmapped linear solve: one batched FP64 adjoint solve
    lambda = J_bc^-1 V (symmetric elimination, lambda[bc] = 0) on the stacked
    problem, then the parameter and traction pullbacks of the stacked
    problem, whose cells and loaded faces are the lanes' in lane order."""
    global _VMAP_FN
    if _VMAP_FN is not None:
        return _VMAP_FN
    import torch
    from .solvers.adjoint import _rebuild, parameter_gradients

    class VmapSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, call, *tensors):
            tp = _rebuild(call["tp"], call["paths"], [t.detach() for t in tensors])
            keep = {}
            U = call["run"](tp, call["bc_vals"], keep=keep)
            ctx.call, ctx.keep = call, keep
            ctx.shapes = [tuple(t.shape) for t in tensors]
            ctx.meta = [(t.dtype, t.device) for t in tensors]
            ctx.save_for_backward(U)
            return U

        @staticmethod
        def backward(ctx, V):
            from .assembler import assemble_csr
            (U,) = ctx.saved_tensors
            k, opts = ctx.keep, ctx.call["opts"]
            bp, ts, bcb, tpd = k["bp"], k["ts"], k["bc"], k["tp"]
            J, dinv = assemble_csr(bp, ts, tpd, None, bcb, True, torch.float64, jacobi_shift=opts.jacobi_shift)
            rhs = V.detach().to(device=ts.device, dtype=torch.float64).reshape(-1).clone()
            if len(bcb.bc_rows):
                rhs.index_fill_(0, bcb.device(ts.device)[0].long(), 0.0)
            lam, _, _, _ = pcg_batched(bp, J, rhs, None, dinv if opts.use_jacobi_preconditioner else None,
                                       opts.tol, opts.atol, opts.maxiter, opts.check_every)
            grads = parameter_gradients(bp, ts, tpd, bcb, U.reshape(-1), lam, ctx.call["paths"], ctx.shapes,
                                        ctx.needs_input_grad[1:])
            return (None, *[g if g is None else g.to(device=d, dtype=dt) for g, (dt, d) in zip(grads, ctx.meta)])

    _VMAP_FN = VmapSolve
    return _VMAP_FN


__all__ = ["BatchedProblem", "batch_bc", "batch_params", "create_batched_solver", "pcg_batched", "BatchedStats",
           "SegKrylovRecord", "create_vmap_solver"]
