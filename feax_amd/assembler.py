"""Device assembly: TracedStructure, CSR Jacobian and residual (feax.assembler API).

feax entry points mirrored (feax/assembler.py):
  get_res                         1618-1667  -> residual kernel, list of (N, 3)
  create_J_bc_csr_parametric      1695-1713  -> fn(sol, tp, bc, ts) -> CSRMatrix
  create_res_bc_parametric        1833-1858  -> fn(sol, tp, bc, ts) -> res_bc
  create_res_J_bc_csr_parametric  1716-1737  -> fn(sol, tp, bc, ts) -> (res_bc, J)
TracedStructure mirrors feax/traced_structure.py:45-164: the mesh-sized
structure arrays, resident on the device, passed to every call.
"""

import ctypes

import numpy as np

from . import _native as nat
from .csr import CSRMatrix
from .materials import SurfaceTraction, TensorMap


class TracedStructure:
    """Device-resident structure of a Problem (SoA coordinates, connectivity,
    node->element incidence, CSR pattern, loaded faces)."""

    def __init__(self, problem, device="cuda"):
        import torch
        dev = torch.device(device)
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        P = problem.points
        self.device = dev
        self.ele_type = problem.ele_type
        self.tag = problem.kernel_tag
        self.coupled = getattr(problem, "coupled", False)
        # components per node of the surface direction table: (u, T) -> 3 + 1
        self.vec = 4 if self.coupled else problem.vec
        self.num_face_quads = problem.fes[0].num_face_quads
        self.num_nodes = problem.num_total_nodes
        self.num_cells = problem.num_cells
        self.n = problem.num_total_dofs_all_vars
        self.nse = problem.csr_nse
        self.x = torch.as_tensor(np.ascontiguousarray(P[:, 0]), **f64)
        self.y = torch.as_tensor(np.ascontiguousarray(P[:, 1]), **f64)
        self.z = torch.as_tensor(np.ascontiguousarray(P[:, 2]), **f64) if P.shape[1] > 2 else None
        self.cells = torch.as_tensor(problem.cells.reshape(-1), **i32)
        self.inc_ptr = torch.as_tensor(problem.node_inc_ptr, **i32)
        self.inc = torch.as_tensor(problem.node_inc, **i32)
        self.rank = torch.as_tensor(problem.elem_rank, dtype=torch.uint8, device=dev)
        self.indptr = torch.as_tensor(problem.csr_indptr, **i32)
        self.indices = torch.as_tensor(problem.csr_indices, **i32)
        self.num_faces = len(problem.all_faces)
        self.faces = torch.as_tensor(problem.all_faces.reshape(-1), **i32) if self.num_faces else None
        self.face_inc_ptr = torch.as_tensor(problem.face_inc_ptr, **i32)
        self.face_inc = torch.as_tensor(problem.face_inc, **i32)
        self.face_offsets = problem.face_offsets
        # Per-face traction direction from the problem's surface maps.
        maps = problem.get_surface_maps() or []
        fdir = np.zeros((self.num_faces, self.vec))
        for i in range(len(problem.boundary_inds_list)):
            if i < len(maps):
                m = maps[i]
                if not isinstance(m, SurfaceTraction):
                    raise NotImplementedError(
                        "surface maps on the HIP path are feax_amd.materials.SurfaceTraction (dead loads)")
                d = np.asarray(m.direction, float)
                if self.coupled:
                    # a SurfaceFlux loads the temperature, a SurfaceTraction the displacement
                    from .materials import SurfaceFlux
                    full = np.zeros(4)
                    if isinstance(m, SurfaceFlux):
                        full[3] = d[0]
                    elif d.shape == (3,):
                        full[:3] = d
                    else:
                        raise ValueError(f"surface map direction {d} must have 3 components here")
                    d = full
                if d.shape != (self.vec,):
                    raise ValueError(f"surface map direction {d} must have {self.vec} components here")
                fdir[self.face_offsets[i]:self.face_offsets[i + 1]] = d
        self.fdir = torch.as_tensor(fdir.reshape(-1), **f64)
        # An optional order for the HEX8 per-node gather (morton_order: each
        # XCD then gathers a compact region and an element's workspace stays
        # in its L2 for all 8 readers; it halves the gather's HBM reads at 1M
        # DOF but measured 6 % slower, the kernel being bound by its index
        # arithmetic).  Results do not depend on it; None = natural order.
        self.node_order = None
        self._cache = {}
        self._bcol = None

    def b3map(self, value_bytes, threads=256):
        """Block-column map (nat.B3Map) of the block-3 pattern for the SpMV of
        the PCG with `value_bytes`-wide matrix values, or None when a
        workgroup's unique column nodes would not fit its LDS staging.  One
        workgroup of `threads` (256 or 512) runs threads / lanes consecutive
        nodes (lanes from fx_csr_block3_lanes); its sorted unique column
        nodes are ucol[uptr[b]:uptr[b + 1]] and lidx gives each 3x3 block the
        position of its column node in that list.  Built once per size."""
        import numpy as np
        import torch
        lanes = int(nat.lib().fx_csr_block3_lanes(self.n, self.nse, int(value_bytes)))
        G = int(threads) // lanes
        cache = self.__dict__.setdefault("_b3maps", {})
        if G in cache:
            return cache[G]
        N = self.n // 3
        indptr = self.indptr.cpu().numpy().astype(np.int64)
        bcol = self.bcol().cpu().numpy().astype(np.int64)
        deg = (indptr[3 * np.arange(N) + 1] - indptr[3 * np.arange(N)]) // 3
        blk = np.repeat(np.arange(N, dtype=np.int64) // G, deg)
        uniq, inv = np.unique(blk * N + bcol, return_inverse=True)
        nblk = (N + G - 1) // G
        uptr = np.zeros(nblk + 1, np.int64)
        np.add.at(uptr, uniq // N + 1, 1)
        uptr = np.cumsum(uptr)
        nu = np.diff(uptr)
        m = None
        if 3 * 8 * int(nu.max()) <= 48 * 1024:
            lidx = (inv.reshape(-1) - uptr[blk]).astype(np.int16)
            keep = (torch.as_tensor(lidx, device=self.device), torch.as_tensor(uptr.astype(np.int32), device=self.device),
                    torch.as_tensor((uniq % N).astype(np.int32), device=self.device))
            m = nat.B3Map(keep, G, int(nu.max()))
        cache[G] = m
        return m

    def b3sym(self, nodes_per_block=64):
        """Upper-block storage map (nat.B3Sym) of the block-3 pattern for the
        symmetric SpMV kernels (fx_pcg_b3sym_*: 64 nodes per workgroup for
        the FP32 difference form, 32 for FP64 values), or None when the
        pattern does not fit them (upper_block_tables).  Built once per
        workgroup size."""
        import torch
        cache = self.__dict__.setdefault("_b3syms", {})
        if nodes_per_block in cache:
            return cache[nodes_per_block]
        cache[nodes_per_block] = None
        if self.vec != 3 or self.n % 3 or self.n == 0:
            return None
        tab = upper_block_tables(self.indptr.cpu().numpy(), self.bcol().cpu().numpy(), nodes_per_block)
        if tab is not None:
            dev = self.device
            t = (torch.as_tensor(tab["code"].view(np.int16), device=dev),
                 torch.as_tensor(tab["uptr"], device=dev), torch.as_tensor(tab["ucol"], device=dev),
                 torch.as_tensor(tab["usup"], device=dev), torch.as_tensor(tab["up_ptr"], device=dev))
            cache[nodes_per_block] = nat.B3Sym(t, tab["nodes_per_block"], tab["max_unique"], tab["max_degree"],
                                               tab["num_upper"])
        return cache[nodes_per_block]

    def bcol(self):
        """Column node of each 3x3 block of the CSR pattern (nse/9 int32)."""
        import torch
        if self._bcol is None:
            b = torch.empty(self.nse // 9, dtype=torch.int32, device=self.device)
            nat.call("fx_csr_block3_index", nat.stream(), self.n, nat.ptr(self.indptr), nat.ptr(self.indices),
                     nat.ptr(b))
            self._bcol = b
        return self._bcol

    @classmethod
    def from_problem(cls, problem, device="cuda", free_scratch=False):
        return cls(problem, device)


def morton_order(points):
    """Node indices sorted along a Z-order (Morton) curve of the coordinates
    (one scale for all axes, 21 bits per axis in 3-D, 31 in 2-D)."""
    P = np.asarray(points, np.float64)
    if len(P) == 0:
        return np.zeros(0, np.int32)
    d = P.shape[1]
    bits = 21 if d == 3 else 31
    lo = P.min(axis=0)
    span = max(float((P.max(axis=0) - lo).max()), 1e-300)
    q = np.minimum(((P - lo) / span * (2 ** bits - 1)).astype(np.int64), 2 ** bits - 1)
    key = np.zeros(len(P), np.int64)
    for b in range(bits):
        for k in range(d):
            key |= ((q[:, k] >> b) & 1) << (d * b + (d - 1 - k))
    return np.argsort(key, kind="stable").astype(np.int32)


def upper_block_tables(indptr, bcol, nodes_per_block=64):
    """Host tables of the upper-block storage (fx_b3sym) of a block-3
    pattern: the symmetric difference-form SpMV reads node n's blocks with
    column c >= n from its own contiguous list (from block up_ptr[n]) and a
    lower block (n, c) as the transpose of the stored block (c, n).

    Returns None when the pattern does not fit the kernels: not symmetric, a
    node without its diagonal block, a node degree above 32, or more than
    1023 unique column nodes per workgroup of `nodes_per_block` nodes.
    Otherwise a dict: code (uint16, 32 per node: entry 32 n + kb for node
    n's block kb in bcol order, (l << 6) | (lower << 5) | r, l = the column
    node's position in the workgroup's sorted unique column list, r = the
    stored block's position in its owner's upper list -- the column node's
    for a lower block, else the row node's; 0xFFFF past the node's degree),
    uptr / ucol (the workgroups' unique column lists), usup (up_ptr of each
    ucol entry), up_ptr (N + 1), num_upper, max_unique, max_degree,
    nodes_per_block."""
    G = int(nodes_per_block)
    ip = np.asarray(indptr, np.int64)
    bc = np.asarray(bcol, np.int64)
    N = (len(ip) - 1) // 3
    if N == 0:
        return None
    bstart = ip[3 * np.arange(N)] // 9
    deg = (ip[3 * np.arange(N) + 1] - ip[3 * np.arange(N)]) // 3
    if deg.max() > 32 or len(bc) != int(deg.sum()):
        return None
    node = np.repeat(np.arange(N, dtype=np.int64), deg)
    lower = bc < node
    if np.count_nonzero(bc == node) != N:
        return None                                   # a node without its diagonal block
    nlow = np.bincount(node[lower], minlength=N)
    up_ptr = np.zeros(N + 1, np.int64)
    up_ptr[1:] = np.cumsum(deg - nlow)
    keys = node * N + bc                              # ascending: nodes, then sorted columns
    want = bc * N + node
    mirror = np.searchsorted(keys, want)
    if np.any(mirror >= len(keys)) or np.any(keys[np.minimum(mirror, len(keys) - 1)] != want):
        return None                                   # pattern not symmetric
    kb = np.arange(len(bc)) - bstart[node]
    r = np.where(lower, mirror - bstart[bc] - nlow[bc], kb - nlow[node])
    blk = node // G
    uniq, inv = np.unique(blk * N + bc, return_inverse=True)
    nblk = (N + G - 1) // G
    uptr = np.zeros(nblk + 1, np.int64)
    np.add.at(uptr, uniq // N + 1, 1)
    uptr = np.cumsum(uptr)
    lidx = inv.reshape(-1) - uptr[blk]
    maxu = int(np.diff(uptr).max())
    if maxu > 1023 or r.max() > 31:
        return None
    code = np.full(32 * N, 0xFFFF, np.uint16)
    code[32 * node + kb] = (lidx << 6) | (lower.astype(np.int64) << 5) | r
    ucol = uniq % N
    return dict(code=code, uptr=uptr.astype(np.int32), ucol=ucol.astype(np.int32),
                usup=up_ptr[ucol].astype(np.int32), up_ptr=up_ptr.astype(np.int32), num_upper=int(up_ptr[-1]),
                max_unique=maxu, max_degree=int(deg.max()), nodes_per_block=G)


def _as_dev(a, ts, dtype=None):
    import torch
    dtype = torch.float64 if dtype is None else dtype
    if isinstance(a, torch.Tensor):
        return a.to(device=ts.device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=ts.device)


def resolve_material(problem, traced_params, ts):
    """(fx_material, param tensor) for the element kernels.

    The parameter is volume_vars[0], classified by its global shape like
    feax/assembler.py:69-103 (node var (N,) or cell var (E,)); a tensor map
    with a fixed modulus and no volume vars gets a constant cell var.
    """
    import torch
    tm = problem.get_tensor_map()
    if tm is None:
        tm = problem.get_energy_density()
    if getattr(problem, "coupled", False):
        tm = problem.get_weak_form()
        if not getattr(tm, "coupled", False):
            raise NotImplementedError("a coupled (u, T) problem's get_weak_form() returns a "
                                      "feax_amd.materials.ThermoElasticCoupled on the HIP path")
    if not isinstance(tm, TensorMap):
        raise NotImplementedError(
            "the HIP element kernels need get_tensor_map()/get_energy_density() to return a "
            "feax_amd.materials law (IsotropicLinearElastic, SIMPLinearElastic, CompressibleNeoHookean)")
    if problem.get_mass_map() is not None:
        raise NotImplementedError("mass maps are not on the HIP path")
    vv = traced_params.volume_vars if traced_params is not None else ()
    if len(vv):
        p = _as_dev(vv[0], ts)
        if p.ndim == 1 and p.shape[0] == ts.num_nodes:
            loc = nat.FX_PARAM_NODE
        elif p.ndim == 1 and p.shape[0] == ts.num_cells:
            loc = nat.FX_PARAM_CELL
        else:
            raise NotImplementedError(f"volume var of shape {tuple(p.shape)} (node (N,) or cell (E,) only)")
    else:
        E = getattr(tm, "fixed_param", getattr(tm, "E", None))
        if E is None:
            raise ValueError("tensor map needs its modulus from TracedParams.volume_vars[0]")
        key = ("constE", E)
        if key not in ts._cache:
            ts._cache[key] = torch.full((ts.num_cells,), float(E), dtype=torch.float64, device=ts.device)
        p, loc = ts._cache[key], nat.FX_PARAM_CELL
    mat = tm.material(loc)
    alpha = getattr(tm, "alpha", None)
    if alpha is not None and not getattr(tm, "coupled", False):
        # thermo-elastic law: the temperature is volume_vars[1], a node var
        if ts.tag != "hex8":
            raise NotImplementedError("thermal strain is built for HEX8 meshes (vec = 3)")
        if len(vv) < 2:
            raise ValueError("a thermo-elastic law takes the temperature as volume_vars[1] (node variable)")
        T = _as_dev(vv[1], ts).to(torch.float64).contiguous()
        if T.ndim != 1 or T.shape[0] != ts.num_nodes:
            raise NotImplementedError(f"the temperature must be a node variable ({ts.num_nodes},)")
        mat.temperature, mat.alpha, mat.t_ref = T.data_ptr(), float(alpha), float(tm.t_ref)
        mat._temperature = T      # keeps the device array alive with the struct
    return mat, p


def thermal_vjp(problem, ts, traced_params, lam, bc=None, scale=1.0):
    """scale * (dR/dT)^T lam for the temperature node variable
    (volume_vars[1]) of a thermo-elastic law (fx_hex8_thermal_vjp_f64)."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    if not mat.temperature:
        raise ValueError("the tensor map has no temperature field")
    lam = _as_dev(lam, ts)
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    grad = torch.empty(ts.num_nodes, dtype=torch.float64, device=ts.device)
    wsb = int(nat.lib().fx_hex8_param_vjp_workspace_bytes(ts.num_cells, nat.FX_PARAM_NODE))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=ts.device)
    nat.call("fx_hex8_thermal_vjp_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat),
             nat.ptr(p), nat.ptr(lam), nat.ptr(mask), float(scale), nat.ptr(grad), nat.ptr(ws), wsb)
    return grad


def surface_values(problem, traced_params, ts):
    """(F*q,) float64: surface_vars[i][0] of every loaded boundary,
    concatenated (q = 4 face Gauss points for HEX8, 1 for TET4)."""
    import torch
    if ts.num_faces == 0:
        return None
    q = ts.num_face_quads
    sv = traced_params.surface_vars if traced_params is not None else []
    parts = []
    for i in range(len(problem.boundary_inds_list)):
        nf = int(ts.face_offsets[i + 1] - ts.face_offsets[i])
        if i < len(sv) and len(sv[i]):
            t = _as_dev(sv[i][0], ts)
            if t.numel() == nf:
                t = t.reshape(nf, 1).expand(nf, q)
            parts.append(t.reshape(nf, q))
        else:
            parts.append(torch.zeros((nf, q), dtype=torch.float64, device=ts.device))
    return torch.cat(parts).reshape(-1).contiguous()


def hex8_only(ts, what):
    """The kernels of `what` exist for HEX8 only."""
    if ts.tag != "hex8":
        raise NotImplementedError(f"{what} is implemented for HEX8 meshes only (got {ts.ele_type})")


# Element workspace of the two-pass assembly / residual (576 or 24 values
# per element), one growing buffer per device; above the cap the kernels
# run single-pass with no workspace.
ELEM_WORKSPACE_CAP = 4 << 30
_ELEM_WS = {}


def elem_workspace(ts, nbytes):
    """(pointer, size) of the element workspace, or (None, 0) above the cap
    (HEX8 then runs single-pass; the TET4 kernels always take it)."""
    import torch
    if nbytes <= 0 or (nbytes > ELEM_WORKSPACE_CAP and ts.tag == "hex8"):
        return None, 0
    key = str(ts.device)
    ws = _ELEM_WS.get(key)
    if ws is None or ws.numel() < nbytes:
        _ELEM_WS.pop(key, None)
        ws = torch.empty(int(nbytes), dtype=torch.uint8, device=ts.device)
        _ELEM_WS[key] = ws
    return ws.data_ptr(), ws.numel()


def upper_assembly_ok(ts, mat, symmetric, ws_bytes):
    """Whether the FP32 difference form can be assembled straight into the
    upper-block storage (FX_ASM_UPPER): a HEX8 vec = 3 mesh whose pattern has
    the upper-block tables, a linear law (its element matrices are exactly
    symmetric, so the eliminated matrix is bitwise symmetric), symmetric
    elimination and an element workspace."""
    return (ts.tag == "hex8" and ts.vec == 3 and bool(symmetric) and ws_bytes > 0
            and mat.kind in (nat.FX_MAT_LINEAR_ELASTIC, nat.FX_MAT_SIMP) and ts.b3sym() is not None)


def assemble_csr(problem, ts, traced_params, sol=None, bc=None, symmetric=True, dtype=None, diff_form=False,
                 jacobi_shift=None, upper=False, compute32=False):
    """BC-eliminated CSR Jacobian on the device (fx_hex8_assemble_* / fx_tet4_assemble_* /
    fx_quad4_assemble_*).

    diff_form: FP32 difference form for the mixed solver.  upper: with
    diff_form, store it as the upper-block storage of a symmetric matrix
    (form "sym3": the operator of the difference form from about half the
    values; used when upper_assembly_ok, else the full difference form is
    returned).  jacobi_shift: also return dinv = 1 / (diag + shift) (FP64)
    computed in the same pass.  compute32 (HEX8, FP32 values only,
    FX_ASM_COMPUTE32): element matrices formed in FP32 arithmetic as well
    (by default they are formed in FP64 and rounded once at the store).
    Returns the CSRMatrix, or (CSRMatrix, dinv) when jacobi_shift is given.
    """
    import torch
    dtype = torch.float64 if dtype is None else dtype
    mat, p = resolve_material(problem, traced_params, ts)
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    u = _as_dev(sol, ts) if sol is not None else None
    dinv = torch.empty(ts.n, dtype=torch.float64, device=ts.device) if jacobi_shift is not None else None
    name = f"fx_{ts.tag}_assemble_f64" if dtype == torch.float64 else f"fx_{ts.tag}_assemble_f32"
    flags = nat.FX_ASM_DIFF3 if diff_form else 0
    if compute32:
        if ts.tag != "hex8" or dtype != torch.float32:
            raise ValueError("compute32 is the HEX8 FP32 assembly's option")
        flags |= nat.FX_ASM_COMPUTE32
    esize = torch.empty(0, dtype=dtype).element_size()
    wsb = getattr(nat.lib(), f"fx_{ts.tag}_assemble_workspace_bytes")(ts.num_cells, esize, flags)
    wp, wn = elem_workspace(ts, wsb)
    sym = None
    if upper and diff_form and dtype == torch.float32 and upper_assembly_ok(ts, mat, symmetric, wn):
        sym = ts.b3sym()
        flags |= nat.FX_ASM_UPPER
    data = torch.empty(9 * sym.num_upper if sym is not None else ts.nse, dtype=dtype, device=ts.device)
    extra = ()
    if ts.tag == "hex8":
        extra = (nat.ptr(ts.node_order), nat.ptr(sym.up_ptr) if sym is not None else None)
    nat.call(name, nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y), nat.ptr(ts.z),
             nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), nat.ptr(ts.rank), nat.ptr(ts.indptr),
             nat.ptr(ts.indices), ctypes.addressof(mat), nat.ptr(p), nat.ptr(u), nat.ptr(mask), int(bool(symmetric)),
             flags, float(jacobi_shift or 0.0), nat.ptr(dinv), nat.ptr(data), wp, wn, *extra)
    b3 = ts.vec == 3
    form = "sym3" if sym is not None else ("diff3" if diff_form else "csr")
    A = CSRMatrix(data, ts.indptr, ts.indices, (ts.n, ts.n), block3=b3, form=form,
                  bcol=ts.bcol() if b3 else None)
    if b3:
        A.b3map = ts.b3map
        A.b3sym = ts.b3sym
        # a linear law with symmetric elimination assembles a bitwise-symmetric
        # matrix: the solvers may run it from upper-block storage
        A.sym_hint = bool(symmetric) and mat.kind in (nat.FX_MAT_LINEAR_ELASTIC, nat.FX_MAT_SIMP)
    return A if jacobi_shift is None else (A, dinv)


def residual(problem, ts, traced_params, sol, bc=None, negate=False, out=None):
    """R(u) (or -R_bc(u) with negate) as a flat float64 device vector."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts)
    res = torch.empty(ts.n, dtype=torch.float64, device=ts.device) if out is None else out
    sv = surface_values(problem, traced_params, ts)
    if bc is not None and len(bc.bc_rows):
        rows, _, vals = bc.device(ts.device)
        nbc = len(bc.bc_rows)
    else:
        rows = vals = None
        nbc = 0
    nf = ts.num_faces
    nat.call(f"fx_{ts.tag}_residual_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat), nat.ptr(p), nat.ptr(u),
             nf, nat.ptr(ts.faces) if nf else None, nat.ptr(ts.face_inc_ptr) if nf else None,
             nat.ptr(ts.face_inc) if nf else None, nat.ptr(sv), nat.ptr(ts.fdir) if nf else None, nbc,
             nat.ptr(rows), nat.ptr(vals), int(bool(negate)), nat.ptr(res),
             *elem_workspace(ts, getattr(nat.lib(), f"fx_{ts.tag}_residual_workspace_bytes")(ts.num_cells)))
    return res


def volume_residual(problem, ts, traced_params, sol, bc=None, negate=False, out=None):
    """The volume part of R_bc: no traction, BC rows u[bc] - 0 when bc is given
    (K_bc u for a linear law with u[bc] = 0; the adjoint operator)."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts)
    res = torch.empty(ts.n, dtype=torch.float64, device=ts.device) if out is None else out
    if bc is not None and len(bc.bc_rows):
        rows = bc.device(ts.device)[0]
        key = ("zero_bc_vals", len(bc.bc_rows))
        if key not in ts._cache:
            ts._cache[key] = torch.zeros(len(bc.bc_rows), dtype=torch.float64, device=ts.device)
        vals, nbc = ts._cache[key], len(bc.bc_rows)
    else:
        rows = vals = None
        nbc = 0
    nat.call(f"fx_{ts.tag}_residual_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat),
             nat.ptr(p), nat.ptr(u), 0, None, None, None, None, None, nbc, nat.ptr(rows), nat.ptr(vals),
             int(bool(negate)), nat.ptr(res), *elem_workspace(ts, getattr(nat.lib(), f"fx_{ts.tag}_residual_workspace_bytes")(ts.num_cells)))
    return res


def jvp(problem, ts, traced_params, sol, v, bc=None, out=None):
    """Matrix-free J_bc(sol) v (fx_hex8_jvp_f64): feax's Krylov operator
    create_matfree_res_J_parametric (feax/assembler.py:1747-1778)."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts) if sol is not None else None
    v = _as_dev(v, ts)
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    res = torch.empty(ts.n, dtype=torch.float64, device=ts.device) if out is None else out
    hex8_only(ts, "the matrix-free operator")
    nat.call("fx_hex8_jvp_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat), nat.ptr(p),
             nat.ptr(u), nat.ptr(v), nat.ptr(mask), nat.ptr(res))
    return res


def create_matfree_res_J_parametric(problem, symmetric=True):
    """fn(sol, tp, bc, ts) -> (res_bc, matvec) with matvec(v) = J_bc(sol) v
    (feax/assembler.py:1747-1778).  Symmetric elimination only."""
    if not symmetric:
        raise NotImplementedError("the matrix-free operator is built for symmetric elimination")

    def fn(sol_flat, traced_params, bc, ts=None):
        ts_ = _ts_or_build(problem, ts)
        res = residual(problem, ts_, traced_params, sol_flat, bc)
        return res, (lambda v: jvp(problem, ts_, traced_params, sol_flat, v, bc))
    return fn


def param_vjp(problem, ts, traced_params, sol, lam, bc=None, scale=1.0):
    """scale * (dR/dp)^T lam for p = volume_vars[0] (fx_hex8_param_vjp_f64 /
    fx_tet4_param_vjp_f64 / fx_quad4_param_vjp_f64);
    lam is read as 0 on the Dirichlet rows of bc.  Returns a float64 device
    tensor shaped like the parameter (cell (E,) or node (N,))."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts)
    lam = _as_dev(lam, ts)
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    grad = torch.empty_like(p, dtype=torch.float64)
    if ts.tag == "hex8ut":
        # E enters only the displacement residual, sigma(grad u, T): the HEX8
        # thermo-elastic pullback on the u part, T = the state's temperature.
        n3 = 3 * ts.num_nodes
        T = u[n3:].contiguous()
        mat.temperature = T.data_ptr()
        u, lam = u[:n3].contiguous(), lam[:n3].contiguous()
        mask = mask[:n3].contiguous() if mask is not None else None
        wsb = int(nat.lib().fx_hex8_param_vjp_workspace_bytes(ts.num_cells, mat.param_loc))
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=ts.device)
        nat.call("fx_hex8_param_vjp_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
                 nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat),
                 nat.ptr(p), nat.ptr(u), nat.ptr(lam), nat.ptr(mask), float(scale), nat.ptr(grad), nat.ptr(ws), wsb)
        return grad
    wsb = int(getattr(nat.lib(), f"fx_{ts.tag}_param_vjp_workspace_bytes")(ts.num_cells, mat.param_loc))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=ts.device)
    nat.call(f"fx_{ts.tag}_param_vjp_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat),
             nat.ptr(p), nat.ptr(u), nat.ptr(lam), nat.ptr(mask), float(scale), nat.ptr(grad), nat.ptr(ws), wsb)
    return grad


def traction_vjp(problem, ts, lam, bc=None, scale=1.0):
    """scale * d(lam . R)/d svar per loaded boundary: a list of (F_i, q)
    float64 device tensors (fx_hex8_traction_vjp_f64, q = 4 /
    fx_tet4_traction_vjp_f64, q = 1 / fx_quad4_traction_vjp_f64, q = 2)."""
    import torch
    if ts.num_faces == 0:
        return []
    lam = _as_dev(lam, ts)
    mask = bc.device(ts.device)[1] if (bc is not None and len(bc.bc_rows)) else None
    q = ts.num_face_quads
    g = torch.empty(ts.num_faces * q, dtype=torch.float64, device=ts.device)
    if ts.tag == "hex8ut":
        # traction on u (HEX8 pullback) plus flux on T (scalar HEX8 pullback)
        n3 = 3 * ts.num_nodes
        fd = ts.fdir.reshape(-1, 4)
        gT = torch.empty_like(g)
        parts = ((fd[:, :3].contiguous(), lam[:n3].contiguous(), None if mask is None else mask[:n3].contiguous(),
                  "fx_hex8_traction_vjp_f64", g),
                 (fd[:, 3].contiguous(), lam[n3:].contiguous(), None if mask is None else mask[n3:].contiguous(),
                  "fx_hex8s_traction_vjp_f64", gT))
        for fdir, lm, mk, name, out in parts:
            nat.call(name, nat.stream(), ts.num_faces, nat.ptr(ts.x), nat.ptr(ts.y), nat.ptr(ts.z), nat.ptr(ts.cells),
                     nat.ptr(ts.faces), nat.ptr(fdir), nat.ptr(lm), nat.ptr(mk), float(scale), nat.ptr(out))
        g = (g + gT).reshape(-1, q)
        off = ts.face_offsets
        return [g[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    nat.call(f"fx_{ts.tag}_traction_vjp_f64", nat.stream(), ts.num_faces, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.faces), nat.ptr(ts.fdir), nat.ptr(lam), nat.ptr(mask),
             float(scale), nat.ptr(g))
    g = g.reshape(-1, q)
    off = ts.face_offsets
    return [g[off[i]:off[i + 1]] for i in range(len(off) - 1)]


def compliance_sensitivity(problem, ts, traced_params, sol):
    """Fused dC/drho_e = -u_e^T dK_e/drho_e u_e for a linear law with a cell
    variable (fx_simp_compliance_sens_f64 and the TET4 / QUAD4 variants): the
    self-adjoint gradient of the compliance when the Dirichlet values are
    zero."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts)
    dc = torch.empty(ts.num_cells, dtype=torch.float64, device=ts.device)
    name = "fx_simp_compliance_sens_f64" if ts.tag == "hex8" else f"fx_{ts.tag}_simp_compliance_sens_f64"
    nat.call(name, nat.stream(), ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y), nat.ptr(ts.z),
             nat.ptr(ts.cells), ctypes.addressof(mat), nat.ptr(p), nat.ptr(u), nat.ptr(dc))
    return dc


def strain_energy(problem, ts, traced_params, sol, with_partial=False):
    """Neo-Hookean strain energy Pi(u) (device scalar) and, with_partial,
    dPi/dmu_e per element (fx_hex8_energy_f64 / fx_tet4_energy_f64)."""
    import torch
    mat, p = resolve_material(problem, traced_params, ts)
    u = _as_dev(sol, ts)
    out = torch.empty(1, dtype=torch.float64, device=ts.device)
    dmu = torch.empty(ts.num_cells, dtype=torch.float64, device=ts.device) if with_partial else None
    if ts.tag not in ("hex8", "tet4"):
        raise NotImplementedError("the strain energy is implemented for HEX8 and TET4 meshes (vec = 3)")
    wsb = int(getattr(nat.lib(), f"fx_{ts.tag}_energy_workspace_bytes")(ts.num_cells))
    ws = torch.empty(wsb, dtype=torch.uint8, device=ts.device)
    nat.call(f"fx_{ts.tag}_energy_f64", nat.stream(), ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y), nat.ptr(ts.z),
             nat.ptr(ts.cells), ctypes.addressof(mat), nat.ptr(p), nat.ptr(u), nat.ptr(out), nat.ptr(dmu),
             nat.ptr(ws), wsb)
    return (out[0], dmu) if with_partial else out[0]


def create_energy_fn(problem, traced_structure=None):
    """energy(u_flat, traced_params) -> Pi = sum_e sum_q psi(grad u, mu) JxW
    (feax/assembler.py:1866-1961) for the Neo-Hookean law, as a 0-d float64
    tensor differentiable through torch.autograd in u (dPi/du = the volume
    residual, fx_hex8_residual_f64) and in a cell variable mu (the partial
    dPi/dmu_e of fx_hex8_energy_f64)."""
    import torch

    class EnergyFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u, mu, tp_static):
            ts = _ts_or_build(problem, traced_structure)
            tp = tp_static if mu is None else _with_volume_var(tp_static, mu.detach())
            need_mu = mu is not None and mu.requires_grad
            Pi, dmu = strain_energy(problem, ts, tp, u.detach(), with_partial=need_mu)
            ctx.save_for_backward(u.detach())
            ctx.tp, ctx.dmu, ctx.ts = tp, dmu, ts
            return Pi

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            gu = gm = None
            if ctx.needs_input_grad[0]:
                gu = g * volume_residual(problem, ctx.ts, ctx.tp, u)
            if ctx.needs_input_grad[1]:
                gm = g * ctx.dmu
            return gu, gm, None

    def energy(u_flat, traced_params):
        u = getattr(u_flat, "dofs", u_flat)
        vv = traced_params.volume_vars
        mu = vv[0] if len(vv) and isinstance(vv[0], torch.Tensor) else None
        return EnergyFunction.apply(u.reshape(-1), mu, traced_params)

    return energy


def _with_volume_var(tp, v):
    from .traced_params import TracedParams
    return TracedParams(volume_vars=(v,) + tuple(tp.volume_vars[1:]), surface_vars=list(tp.surface_vars))


def _ts_or_build(problem, ts):
    if ts is not None:
        return ts
    cached = getattr(problem, "_default_ts", None)
    if cached is None:
        cached = TracedStructure.from_problem(problem)
        problem._default_ts = cached
    return cached


def get_res(problem, sol_list, traced_params, ts=None):
    """Residual as a list of (N, vec) arrays (feax/assembler.py:1618-1667)."""
    ts = _ts_or_build(problem, ts)
    flat = sol_list[0].reshape(-1) if isinstance(sol_list, (list, tuple)) else sol_list
    return problem.unflatten_fn_sol_list(residual(problem, ts, traced_params, flat))


def get_jacobian(problem, sol_list, traced_params, ts=None):
    """The assembled tangent without Dirichlet conditions as a CSRMatrix in
    the problem's view (feax/assembler.py:1517-1566)."""
    import torch
    ts = _ts_or_build(problem, ts)
    flat = sol_list[0].reshape(-1) if isinstance(sol_list, (list, tuple)) else sol_list
    return view_csr(problem, ts, assemble_csr(problem, ts, traced_params, flat, None, True, torch.float64))


def get_jacobian_info(problem, sol_list, traced_params):
    """nnz, shape and matrix_view of the Jacobian without assembling it
    (feax/assembler.py:1568-1616)."""
    n = problem.num_total_dofs_all_vars
    nnz = len(problem.view_pattern()[1]) if getattr(problem, "matrix_view", None) is not None and \
        problem.matrix_view.name != "FULL" else int(problem.csr_nse)
    return {"nnz": int(nnz), "shape": (n, n), "matrix_view": getattr(problem, "matrix_view", None)}


def create_J_bc_csr_function(problem, bc, symmetric=True, traced_structure=None):
    """(sol_flat, traced_params) -> the BC-applied Jacobian as a CSRMatrix
    (feax/assembler.py:1670-1693)."""
    import torch

    def J_bc_func(sol_flat, traced_params):
        ts = _ts_or_build(problem, traced_structure)
        A = assemble_csr(problem, ts, traced_params, sol_flat, bc, symmetric, torch.float64)
        return view_csr(problem, ts, A)
    return J_bc_func


def create_res_bc_function(problem, bc, traced_structure=None):
    """(sol_flat, traced_params) -> the residual with the Dirichlet rows set
    to u - bc_vals (feax/assembler.py:1797-1830)."""
    def res_bc_func(sol_flat, traced_params):
        ts = _ts_or_build(problem, traced_structure)
        return residual(problem, ts, traced_params, sol_flat, bc)
    return res_bc_func


def view_csr(problem, ts, A):
    """The problem's matrix_view of a FULL CSRMatrix: the UPPER (j >= i) or
    LOWER (j <= i) slots of the pattern, i.e. what feax assembles for that
    view (the dedupe of the filtered (I, J), feax/problem.py:218-231) --
    feax's BC elimination acts slot by slot, so filtering commutes with it.
    FULL returns A itself."""
    import torch
    from .problem import MatrixView
    if problem.matrix_view == MatrixView.FULL:
        return A
    key = ("view", problem.matrix_view)
    if key not in ts._cache:
        indptr, indices, slots = problem.view_pattern()
        dev = ts.device
        ts._cache[key] = (torch.as_tensor(np.ascontiguousarray(indptr, np.int32), device=dev),
                          torch.as_tensor(np.ascontiguousarray(indices, np.int32), device=dev),
                          torch.as_tensor(np.ascontiguousarray(slots, np.int64), device=dev))
    indptr, indices, slots = ts._cache[key]
    return CSRMatrix(A.data.index_select(0, slots), indptr, indices, A.shape, block3=False)


def create_J_bc_csr_parametric(problem, symmetric=True, dtype=None):
    """fn(sol, traced_params, bc, ts) -> J_bc as a CSRMatrix in the
    problem's matrix_view (feax/assembler.py:1695-1713).  The solvers of
    this package always work on the FULL matrix."""
    def fn(sol_flat, traced_params, bc, ts=None):
        ts = _ts_or_build(problem, ts)
        return view_csr(problem, ts, assemble_csr(problem, ts, traced_params, sol_flat, bc, symmetric, dtype))
    return fn


def create_res_bc_parametric(problem):
    def fn(sol_flat, traced_params, bc, ts=None):
        return residual(problem, _ts_or_build(problem, ts), traced_params, sol_flat, bc)
    return fn


def create_res_J_bc_csr_parametric(problem, symmetric=True):
    jf = create_J_bc_csr_parametric(problem, symmetric)
    rf = create_res_bc_parametric(problem)

    def fn(sol_flat, traced_params, bc, ts=None):
        return rf(sol_flat, traced_params, bc, ts), jf(sol_flat, traced_params, bc, ts)
    return fn
