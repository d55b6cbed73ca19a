"""Geometric-multigrid preconditioned CG on structured HEX8 boxes.

``KrylovSolverOptions(preconditioner=GeometricMultigrid())`` selects it in
create_solver / create_linear_solver.  feax hands ``options.preconditioner``
to jax cg as ``M`` (feax/solvers/common.py:221-239) and ships a matrix-free
multigrid CG for structured bands, NarrowBandCMG (feax/solvers/cmg.py:
95-300).  This is that solver as a preconditioner of the ordinary solve:

* levels: the mesh must be a full StructuredGrid box (feax/spgrid.py:
  28-41); each coarser grid halves every direction whose cell count is
  even (the others are kept) or, when none is even, ceil-halves the odd
  ones, until a level has at most ``coarse_dofs`` DOFs.
* coarse operators: the HEX8 assembly kernel on the coarse grid
  (rediscretisation) with the coarse parameter = node value at the same
  node (node var) or mean over the child cells (cell var; the density
  averaging of cmg.py:375-386, 426-431), the linearisation point injected,
  and the constrained DOFs injected from the fine BC mask.
* smoother: Chebyshev of ``smooth_degree`` in D^-1 A on
  [lmax / cheb_ratio, 1.1 lmax], lmax from ``lmax_iters`` Lanczos steps.
* coarsest: dense inverse (rocSOLVER Cholesky) of the Galerkin product
  P'^T A P' of the level above (fx_gmg_galerkin_dense_f64; at most
  ``max_dense_dofs`` DOFs), else Chebyshev of ``coarse_degree`` on the
  rediscretised coarsest operator over [lmin, 1.1 lmax] (both Lanczos
  estimates).
* KrylovSolverOptions.precision "fp64" runs the V-cycle in FP64, "mixed"
  in FP32 (FP32 copies of every level); the CG is FP64 either way and stops
  on the jax cg rule ||r||^2 <= max(tol^2 ||b||^2, atol^2).

Everything runs on the device (fx_gmg_*, fx_pcg_gmg_f64); the hierarchy's
structure is built once per problem, its values once per operator.
"""

import ctypes
from dataclasses import dataclass

import numpy as np

from .. import _native as nat


@dataclass(frozen=True)
class GeometricMultigrid:
    smooth_degree: int = 2
    cheb_ratio: float = 12.0
    coarse: str = "dense"
    coarse_degree: int = 40
    max_dense_dofs: int = 12000
    max_levels: int = 12
    min_cells: int = 2
    coarse_dofs: int = 1500
    lmax_iters: int = 10
    check_every: int = 4

    def __post_init__(self):
        if self.coarse not in ("dense", "chebyshev"):
            raise ValueError(f"Invalid coarse solver: {self.coarse}. Choose from ('dense', 'chebyshev')")
        if self.smooth_degree < 1 or self.coarse_degree < 1 or self.lmax_iters < 1:
            raise ValueError("smooth_degree, coarse_degree and lmax_iters must be >= 1")
        if not self.cheb_ratio > 1.0:
            raise ValueError("cheb_ratio must be > 1")
        if self.max_levels < 1 or self.min_cells < 1 or self.check_every < 1 or self.coarse_dofs < 1:
            raise ValueError("max_levels, min_cells, coarse_dofs and check_every must be >= 1")


def structured_dims(problem):
    """(nx, ny, nz) of a Problem whose mesh is a full StructuredGrid box in
    its numbering (nodes sorted by (x, y, z) coordinate index, cells in
    grid order with the HEX8 corner order); raises otherwise."""
    cached = problem.__dict__.get("_gmg_grid")
    if cached is not None:
        return cached
    from ..mesh import StructuredGrid
    P = problem.points
    axes = [np.unique(P[:, d]) for d in range(3)]
    dims = tuple(len(a) - 1 for a in axes)
    ok = min(dims) >= 1 and len(P) == np.prod([d + 1 for d in dims]) and problem.num_cells == np.prod(dims)
    if ok:
        grid = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
        ok = np.array_equal(grid, P)
    if ok:
        cells = StructuredGrid(dims).cell_to_nodes(np.arange(problem.num_cells))
        ok = np.array_equal(cells, problem.cells)
    if not ok:
        raise NotImplementedError("GeometricMultigrid needs a full structured box mesh "
                                  "(StructuredGrid(...).to_mesh() numbering)")
    problem.__dict__["_gmg_grid"] = (dims, axes)
    return dims, axes


def coarse_dims(dims, min_cells):
    """Next coarser grid: halve every direction with an even cell count that
    keeps >= min_cells cells; when none is even, ceil-halve the odd ones
    (the last coarse cell then reaches one fine cell past the grid, as
    feax's NarrowBandCMG coarsens, feax/solvers/cmg.py:82-87, 178-182).
    None when no direction can be coarsened."""
    even = tuple(d // 2 if (d % 2 == 0 and d // 2 >= min_cells) else d for d in dims)
    if even != tuple(dims):
        return even
    odd = tuple((d + 1) // 2 if (d >= 3 and (d + 1) // 2 >= min_cells) else d for d in dims)
    return None if odd == tuple(dims) else odd


def _coarse_axis(a, nf, nc):
    """Node coordinates of a coarse direction (every other fine node; one
    extrapolated node past the end of an odd fine direction)."""
    if nc == nf:
        return a
    c = a[::2]
    if len(c) < nc + 1:
        c = np.append(c, a[-1] + (a[-1] - a[-2]))
    return c


def _law(problem):
    tm = problem.get_tensor_map()
    return tm if tm is not None else problem.get_energy_density()


class GMGStructure:
    """Grids, coarse Problems and TracedStructures of the hierarchy (built
    once per problem and option set)."""

    def __init__(self, problem, ts, mg):
        from ..assembler import TracedStructure
        from ..mesh import Mesh, StructuredGrid
        from ..problem import Problem

        class LevelProblem(Problem):
            law = None

            def get_tensor_map(self):
                return self.law

        dims, axes = structured_dims(problem)
        self.dims = [dims]
        self.ts = [ts]
        self.problems = [problem]
        self.inject = []        # host: coarse node -> fine node of the previous level
        while len(self.dims) < mg.max_levels and 3 * int(np.prod([d + 1 for d in self.dims[-1]])) > mg.coarse_dofs:
            c = coarse_dims(self.dims[-1], mg.min_cells)
            if c is None:
                break
            f = [1 if a == b else 2 for a, b in zip(self.dims[-1], c)]
            axes = [_coarse_axis(a, nf, nc) for a, nf, nc in zip(axes, self.dims[-1], c)]
            pts = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
            cells = StructuredGrid(c).cell_to_nodes(np.arange(int(np.prod(c))))
            cp = LevelProblem(Mesh(pts, cells), vec=3, dim=3)
            cp.law = _law(problem)
            nf = self.dims[-1]
            i, j, k = np.meshgrid(*[np.arange(n + 1) for n in c], indexing="ij")
            I, J, K = (np.minimum(fk * ik, n) for fk, ik, n in zip(f, (i, j, k), nf))
            self.inject.append(((I * (nf[1] + 1) + J) * (nf[2] + 1) + K).reshape(-1))
            self.dims.append(c)
            self.problems.append(cp)
            self.ts.append(TracedStructure.from_problem(cp, ts.device))
        self._bcs = {}

    @property
    def nlev(self):
        return len(self.dims)

    def coarse_bcs(self, bc):
        """DirichletBC per coarse level: DOF (node, c) constrained iff the
        fine DOF at the same node is (values 0: corrections only)."""
        from ..DCboundary import DirichletBC
        key = id(bc)
        hit = self._bcs.get(key)
        if hit is not None and hit[0] is bc:
            return hit[1]
        mask = np.asarray(bc.bc_mask, bool) if bc is not None else np.zeros(self.ts[0].n, bool)
        out = [bc]
        for inj in self.inject:
            mask = mask.reshape(-1, 3)[inj].reshape(-1)
            rows = np.nonzero(mask)[0].astype(np.int32)
            out.append(DirichletBC(rows, mask.copy(), np.zeros(len(rows)), len(mask)))
        self._bcs = {key: (bc, out)}
        return out


def structure_for(problem, ts, mg):
    key = (mg.max_levels, mg.min_cells, mg.coarse_dofs, str(ts.device))
    cache = problem.__dict__.setdefault("_gmg_structure", {})
    st = cache.get(key)
    if st is None or st.ts[0] is not ts:
        st = GMGStructure(problem, ts, mg)
        cache[key] = st
    return st


_WS = {}


def _workspace(nbytes, device):
    import torch
    key = str(device)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def _dense_inverse(A, dev):
    """A^-1 (FP64, dense) of the coarsest SPD operator: the CSR scattered to
    a dense matrix on the device, Cholesky factor and inverse from
    rocSOLVER (torch.linalg)."""
    import torch
    n = A.shape[0]
    dense = torch.empty((n, n), dtype=torch.float64, device=dev)
    nat.call("fx_csr_to_dense_f64", nat.stream(), n, nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(A.data),
             nat.ptr(dense))
    return torch.cholesky_inverse(torch.linalg.cholesky(dense))


class GMGHierarchy:
    """The values of one hierarchy: per-level CSR (FP64, and FP32 copies for
    an FP32 V-cycle), dinv, masks, lmax, vectors, dense coarsest inverse."""

    def __init__(self, problem, ts, tp, u0, bc, A, dinv, mg, fp32=False, jacobi_shift=1e-12):
        import torch
        from ..assembler import assemble_csr, resolve_material
        from ..traced_params import TracedParams
        self.mg = mg
        self.fp32 = bool(fp32)
        S = structure_for(problem, ts, mg)
        self.S = S
        dev = ts.device
        stream = nat.stream()
        bcs = S.coarse_bcs(bc)
        mat, p = resolve_material(problem, tp, ts)
        loc = mat.param_loc
        u = u0 if u0 is not None else torch.zeros(ts.n, dtype=torch.float64, device=dev)
        vdt = torch.float32 if self.fp32 else torch.float64
        self._keep = []
        self.A64 = [A]
        self.dinv64 = [dinv]
        structs = []
        last = S.nlev - 1
        dense_last = mg.coarse == "dense" and S.ts[last].n <= mg.max_dense_dofs
        for l in range(S.nlev):
            cts = S.ts[l]
            mask = bcs[l].device(dev)[1] if (bcs[l] is not None and len(bcs[l].bc_rows)) else None
            if l == last and l > 0 and dense_last:
                # Dense coarsest: only its right-hand side, correction and
                # mask are used (the operator is the Galerkin product below).
                x = torch.empty(cts.n, dtype=vdt, device=dev)
                b = torch.empty(cts.n, dtype=vdt, device=dev)
                self._keep += [mask, x, b]
                nx, ny, nz = S.dims[l]
                structs.append(nat.FxGmgLevel(nx, ny, nz, None, None, None, None, nat.ptr(mask), None, 0.0, 0.0,
                                              nat.ptr(x), nat.ptr(b), None, None))
                continue
            if l > 0:
                prev = S.dims[l - 1]
                pc = torch.empty(cts.num_nodes if loc == nat.FX_PARAM_NODE else cts.num_cells, dtype=torch.float64,
                                 device=dev)
                nat.call("fx_gmg_coarsen_f64", stream, loc, 1, *prev, *S.dims[l], nat.ptr(p), nat.ptr(pc))
                uc = torch.empty(cts.n, dtype=torch.float64, device=dev)
                nat.call("fx_gmg_coarsen_f64", stream, nat.FX_PARAM_NODE, 3, *prev, *S.dims[l], nat.ptr(u),
                         nat.ptr(uc))
                Al, dl = assemble_csr(S.problems[l], cts, TracedParams(volume_vars=(pc,)), uc, bcs[l], True,
                                      torch.float64, jacobi_shift=jacobi_shift)
                self.A64.append(Al)
                self.dinv64.append(dl)
                p, u = pc, uc
            else:
                Al, dl = A, dinv
            data, dv = Al.data, dl
            if self.fp32:
                data = torch.empty(cts.nse, dtype=torch.float32, device=dev)
                nat.call("fx_convert_f64_to_f32", stream, cts.nse, 1.0, nat.ptr(Al.data), nat.ptr(data))
                dv = torch.empty(cts.n, dtype=torch.float32, device=dev)
                nat.call("fx_convert_f64_to_f32", stream, cts.n, 1.0, nat.ptr(dl), nat.ptr(dv))
            need_xb = l > 0 or self.fp32
            x = torch.empty(cts.n, dtype=vdt, device=dev) if need_xb else None
            b = torch.empty(cts.n, dtype=vdt, device=dev) if need_xb else None
            r = torch.empty(cts.n, dtype=vdt, device=dev)
            d = torch.empty(cts.n, dtype=vdt, device=dev)
            self._keep += [Al, dl, data, dv, mask, x, b, r, d]
            nx, ny, nz = S.dims[l]
            structs.append(nat.FxGmgLevel(nx, ny, nz, nat.ptr(Al.indptr), nat.ptr(Al.bcol), nat.ptr(data),
                                          nat.ptr(dv), nat.ptr(mask), None, 0.0, 0.0, nat.ptr(x), nat.ptr(b),
                                          nat.ptr(r), nat.ptr(d)))
        self.levels = (nat.FxGmgLevel * S.nlev)(*structs)
        flags = nat.FX_GMG_F32 if self.fp32 else 0
        self.flags = flags
        # Coarsest level: dense inverse when small enough, of the Galerkin
        # product P'^T A P' of the level above (exact across a ceil-halved
        # direction, where rediscretisation is not).
        nc = S.ts[last].n
        self.dense = None
        if dense_last:
            if last > 0:
                Af = self.A64[last - 1]
                mf = bcs[last - 1].device(dev)[1] if len(bcs[last - 1].bc_rows) else None
                mc = bcs[last].device(dev)[1] if len(bcs[last].bc_rows) else None
                Ac = torch.empty((nc, nc), dtype=torch.float64, device=dev)
                ws = _workspace(nat.lib().fx_gmg_galerkin_workspace_bytes(*S.dims[last - 1], *S.dims[last]), dev)
                nat.call("fx_gmg_galerkin_dense_f64", stream, *S.dims[last - 1], *S.dims[last], nat.ptr(Af.indptr),
                         nat.ptr(Af.indices), nat.ptr(Af.data), nat.ptr(mf), nat.ptr(mc), nat.ptr(Ac), nat.ptr(ws),
                         ws.numel())
                inv = torch.cholesky_inverse(torch.linalg.cholesky(Ac))
            else:
                inv = _dense_inverse(self.A64[last], dev)
            if self.fp32:
                inv32 = torch.empty((nc, nc), dtype=torch.float32, device=dev)
                nat.call("fx_convert_f64_to_f32", stream, nc * nc, 1.0, nat.ptr(inv), nat.ptr(inv32))
                inv = inv32
            self.dense = inv
            self.levels[last].dense_inv = nat.ptr(inv)
        # Spectrum bounds of D^-1 A of every level the V-cycle smooths, from
        # the FP64 operators.
        self.lmax, self.lmin = [], []
        for l in range(S.nlev):
            if l == last and self.dense is not None:
                self.lmax.append(0.0)
                self.lmin.append(0.0)
                continue
            Al, dl, nn = self.A64[l], self.dinv64[l], S.ts[l].num_nodes
            ws = _workspace(nat.lib().fx_gmg_eig_bounds_workspace_bytes(nn, mg.lmax_iters), dev)
            lo, hi = ctypes.c_double(0.0), ctypes.c_double(0.0)
            nat.call("fx_gmg_eig_bounds_f64", stream, nn, nat.ptr(Al.indptr), nat.ptr(Al.bcol), nat.ptr(Al.data),
                     nat.ptr(dl), mg.lmax_iters, ctypes.byref(lo), ctypes.byref(hi), nat.ptr(ws), ws.numel())
            self.levels[l].lmax = hi.value
            self.levels[l].lmin = lo.value if l == last else 0.0
            self.lmax.append(hi.value)
            self.lmin.append(lo.value)

    @property
    def nlev(self):
        return self.S.nlev

    def vcycle(self, r, z=None):
        """z = V(r) (FP64 device vectors)."""
        import torch
        z = torch.empty_like(r) if z is None else z
        mg = self.mg
        nat.call("fx_gmg_vcycle", nat.stream(), self.flags, self.nlev, self.levels, mg.smooth_degree,
                 mg.coarse_degree, float(mg.cheb_ratio), nat.ptr(r), nat.ptr(z))
        return z

    def solve(self, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, check_every=None, out=None, graph=True):
        """x = A^-1 b by V-cycle preconditioned CG; returns (x, PCGInfo)."""
        import torch
        from .common import PCGInfo
        n = b.numel()
        x = torch.empty_like(b) if out is None else out
        ws = _workspace(nat.lib().fx_pcg_gmg_workspace_bytes(n), b.device)
        iters = ctypes.c_int32(0)
        rs = ctypes.c_double(0.0)
        mg = self.mg
        flags = self.flags | (0 if graph else nat.FX_PCG_NO_GRAPH)
        nat.call("fx_pcg_gmg_f64", nat.stream(), flags, self.nlev, self.levels, mg.smooth_degree, mg.coarse_degree,
                 float(mg.cheb_ratio), nat.ptr(self.A64[0].data), nat.ptr(b), nat.ptr(x0), float(tol), float(atol),
                 int(maxiter), int(check_every or mg.check_every), nat.ptr(x), ctypes.byref(iters),
                 ctypes.byref(rs), nat.ptr(ws), ws.numel())
        return x, PCGInfo(iters.value, rs.value)


__all__ = ["GeometricMultigrid", "GMGHierarchy", "GMGStructure", "structured_dims"]
