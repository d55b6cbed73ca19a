"""Problem: FE structure for HEX8 / HEX20 / TET4 (vec = 3), QUAD4
(vec = 2), scalar HEX8 / TET4 / QUAD4 (vec = 1) and the coupled (u, T)
HEX8 problem (the feax.Problem API).

Mirrors feax.Problem (feax/problem.py:37-846): constructor arguments,
``location_fns`` / ``boundary_inds_list``, ``matrix_view``, the physics
hooks (get_tensor_map, get_energy_density, get_surface_maps, get_mass_map,
custom_init) and the CSR structure attributes (csr_nse, csr_indptr,
csr_indices, csr_row_of_slot, csr_diag_slots, I / J).

Difference by design: the CSR pattern is built by the native host routine
fx_hex8_pattern from node adjacency (O(E) memory) instead of an argsort
over 576(E+F) int64 keys (feax/problem.py:386-423, ~10 GB at 1M DOF); the
integer arrays are the same.  The nnz-sized raw COO arrays (I, J) are only
materialised if asked for.
"""

import ctypes
from dataclasses import dataclass
from enum import Enum
from typing import Any, Callable, List, Optional, Tuple, Union

import numpy as np

from . import _native as nat
from .fe import boundary_conditions_inds, element_info, physical_quad_points, physical_surface_quad_points
from .mesh import Mesh


class MatrixView(Enum):
    """Which entries the assembled matrix keeps (feax/problem.py:22-35)."""
    FULL = 0
    UPPER = 1
    LOWER = 2


class FiniteElement:
    """The per-variable facts feax.FiniteElement exposes that callers read
    (feax/fe.py:20-94): counts only; geometry lives in the kernels."""

    def __init__(self, mesh, vec, dim, ele_type="HEX8"):
        info = element_info(ele_type)
        self.mesh = mesh
        self.points = mesh.points
        self.cells = mesh.cells
        self.vec = vec
        self.dim = dim
        self.ele_type = ele_type
        self.num_cells = len(mesh.cells)
        self.num_total_nodes = len(mesh.points)
        self.num_total_dofs = self.num_total_nodes * vec
        self.num_nodes = info["npe"]
        self.num_quads = info["num_quads"]
        self.num_face_quads = info["num_face_quads"]
        self.num_faces = len(info["faces"])


@dataclass
class Problem:
    """Vector problem on an explicit mesh: HEX8 / HEX20 / TET4 (vec = dim =
    3), QUAD4 (vec = dim = 2), the scalar (vec = 1) problems, or the coupled
    two-variable HEX8 problem (mesh = [m, m], vec = [3, 1])."""
    coupled = False           # not a dataclass field: set by __post_init__
    mesh: Union[Mesh, List[Mesh]]
    vec: Union[int, List[int]]
    dim: int
    ele_type: Union[str, List[str]] = "HEX8"
    gauss_order: Optional[Union[int, List[int]]] = None
    location_fns: Optional[List[Callable]] = None
    matrix_view: Union[MatrixView, str] = MatrixView.FULL
    additional_info: Tuple[Any, ...] = ()

    def __post_init__(self):
        if isinstance(self.matrix_view, str):
            try:
                self.matrix_view = MatrixView[self.matrix_view.upper()]
            except KeyError:
                raise ValueError(f"Invalid matrix_view: {self.matrix_view}. Must be 'FULL', 'UPPER', or 'LOWER'")
        self.coupled = False
        if isinstance(self.mesh, list):
            if len(self.mesh) == 2:
                self._init_coupled()
                return
            if len(self.mesh) != 1:
                raise NotImplementedError("multi-variable problems on the HIP path: two variables, (u, T) "
                                          "with vec=[3, 1] on one HEX8 mesh")
            self.mesh, self.vec = self.mesh[0], self.vec[0]
            self.ele_type = self.ele_type[0] if isinstance(self.ele_type, list) else self.ele_type
            self.gauss_order = self.gauss_order[0] if isinstance(self.gauss_order, list) else self.gauss_order
        ok = ((self.ele_type in ("HEX8", "HEX20", "TET4") and self.vec == 3 and self.dim == 3)
              or (self.ele_type == "QUAD4" and self.vec == 2 and self.dim == 2)
              or (self.ele_type == "HEX8" and self.vec == 1 and self.dim == 3)
              or (self.ele_type == "QUAD4" and self.vec == 1 and self.dim == 2)
              or (self.ele_type == "TET4" and self.vec == 1 and self.dim == 3))
        if not ok:
            raise NotImplementedError(
                f"the HIP path covers HEX8, HEX20 and TET4 with vec=3, dim=3, QUAD4 with vec=2, dim=2 and scalar HEX8 "
                f"(vec=1, dim=3) (got {self.ele_type}, vec={self.vec}, dim={self.dim})")
        default_order = 0 if self.ele_type == "TET4" else 2
        if self.gauss_order not in (None, default_order):
            raise NotImplementedError(f"{self.ele_type} uses feax's default rule (gauss_order={default_order})")
        self.npe = element_info(self.ele_type)["npe"]
        self.num_vars = 1
        self.fes = [FiniteElement(self.mesh, self.vec, self.dim, self.ele_type)]
        self.points = np.ascontiguousarray(self.mesh.points, dtype=np.float64)
        self.cells = np.ascontiguousarray(self.mesh.cells, dtype=np.int32)
        if self.cells.ndim != 2 or self.cells.shape[1] != self.npe:
            raise ValueError(f"{self.ele_type} cells must be (E, {self.npe}), got {self.cells.shape}")
        self.cells_list = [self.cells]
        self.num_cells = len(self.cells)
        self.num_total_nodes = len(self.points)
        self.num_total_dofs_all_vars = self.vec * self.num_total_nodes
        self.offset = [0]
        self.boundary_inds_list = boundary_conditions_inds(self.points, self.cells, self.location_fns, self.ele_type)
        self._build_csr_assembly_structure()
        self.custom_init(*self.additional_info)

    def _init_coupled(self):
        """Problem(mesh=[m, m], vec=[3, 1], dim=3, ele_type="HEX8"): a
        displacement (variable 0) and a temperature (variable 1) on one HEX8
        mesh in one system, DOFs u 3 n + i then T 3 N + n (feax offsets,
        feax/problem.py:155-212); the physics is get_weak_form() (a
        feax_amd.materials.ThermoElasticCoupled)."""
        m0, m1 = self.mesh
        vec = list(self.vec) if isinstance(self.vec, (list, tuple)) else None
        ets = self.ele_type if isinstance(self.ele_type, list) else [self.ele_type] * 2
        same = (m0 is m1) or (np.array_equal(np.asarray(m0.points), np.asarray(m1.points))
                              and np.array_equal(np.asarray(m0.cells), np.asarray(m1.cells)))
        if vec != [3, 1] or ets != ["HEX8", "HEX8"] or self.dim != 3 or not same:
            raise NotImplementedError("multi-variable problems on the HIP path: (u, T) with vec=[3, 1], dim=3, "
                                      f"HEX8 on one mesh (got vec={self.vec}, ele_type={self.ele_type})")
        if self.gauss_order is not None and any(g not in (None, 2) for g in np.atleast_1d(self.gauss_order)):
            raise NotImplementedError("HEX8 uses feax's default rule (gauss_order=2)")
        if self.matrix_view != MatrixView.FULL:
            raise NotImplementedError("the coupled (u, T) problem assembles the FULL matrix")
        self.coupled = True
        self.mesh_list = [m0, m1]
        self.mesh = m0
        self.vec = [3, 1]
        self.ele_type = "HEX8"
        self.npe = 8
        self.num_vars = 2
        self.fes = [FiniteElement(m0, 3, 3, "HEX8"), FiniteElement(m1, 1, 3, "HEX8")]
        self.points = np.ascontiguousarray(m0.points, dtype=np.float64)
        self.cells = np.ascontiguousarray(m0.cells, dtype=np.int32)
        if self.cells.ndim != 2 or self.cells.shape[1] != 8:
            raise ValueError(f"HEX8 cells must be (E, 8), got {self.cells.shape}")
        self.cells_list = [self.cells, self.cells]
        self.num_cells = len(self.cells)
        self.num_total_nodes = len(self.points)
        self.num_total_dofs_all_vars = 4 * self.num_total_nodes
        self.offset = [0, 3 * self.num_total_nodes]
        self.boundary_inds_list = boundary_conditions_inds(self.points, self.cells, self.location_fns, "HEX8")
        self._build_csr_assembly_structure()
        self.custom_init(*self.additional_info)

    @property
    def kernel_tag(self):
        """Name of the native kernel family: hex8, hex20, tet4, quad4, hex8s /
        tet4s / quad4s (the scalar, vec = 1, problems), or hex8ut (the
        coupled (u, T) problem)."""
        if self.coupled:
            return "hex8ut"
        tag = self.ele_type.lower()
        return tag + "s" if self.vec == 1 else tag

    @property
    def solution_layout(self):
        """(num_nodes, vec) per variable (feax/solution.py:42-90)."""
        return tuple((fe.num_total_nodes, fe.vec) for fe in self.fes)

    @property
    def physical_quad_points(self):
        """(E, num_quads, dim) coordinates of the volume quadrature points
        (feax/problem.py:276-277)."""
        return physical_quad_points(self.points, self.cells, self.ele_type)

    @property
    def physical_surface_quad_points(self):
        """One (F_i, num_face_quads, dim) array per location_fn
        (feax/problem.py:282-308)."""
        return [physical_surface_quad_points(self.points, self.cells, b, self.ele_type)
                for b in self.boundary_inds_list]

    # ------------------------------------------------------------------ structure
    def _build_csr_assembly_structure(self):
        """CSR pattern + assembly maps (feax/problem.py:316-484) via fx_hex8_pattern / fx_tet4_pattern /
        fx_quad4_pattern."""
        N, E, npe = self.num_total_nodes, self.num_cells, self.npe
        n = self.num_total_dofs_all_vars
        L = nat.lib()
        tag = self.kernel_tag
        pattern = getattr(L, f"fx_{tag}_pattern")
        nse = ctypes.c_int64(0)
        maxdeg = ctypes.c_int32(0)
        indptr = np.zeros(n + 1, dtype=np.int32)
        nat.check(pattern(N, E, nat.ptr(self.cells), ctypes.addressof(nse), ctypes.addressof(maxdeg),
                          nat.ptr(indptr), None, None, None, None), f"fx_{tag}_pattern")
        indices = np.empty(nse.value, dtype=np.int32)
        rank = np.empty(E * npe * npe, dtype=np.uint8)
        inc_ptr = np.empty(N + 1, dtype=np.int32)
        inc = np.empty(E * npe, dtype=np.int32)
        nat.check(pattern(N, E, nat.ptr(self.cells), ctypes.addressof(nse), ctypes.addressof(maxdeg),
                          nat.ptr(indptr), nat.ptr(indices), nat.ptr(rank), nat.ptr(inc_ptr), nat.ptr(inc)),
                  f"fx_{tag}_pattern")
        self.csr_nse = int(nse.value)
        self.csr_indptr = indptr
        self.csr_indices = indices
        self.csr_max_degree = int(maxdeg.value)
        self.elem_rank = rank
        self.node_inc_ptr = inc_ptr
        self.node_inc = inc
        self.csr_structurally_symmetric = True   # node adjacency is symmetric
        # Loaded faces of all boundaries, concatenated in location_fns order
        # (the volume-then-boundaries order of feax/problem.py:195-209).
        faces = [b for b in self.boundary_inds_list]
        self.face_offsets = np.cumsum([0] + [len(b) for b in faces])
        allf = (np.concatenate(faces) if faces else np.zeros((0, 2), np.int64)).astype(np.int32)
        self.all_faces = np.ascontiguousarray(allf)
        finc_ptr = np.zeros(N + 1, dtype=np.int32)
        # nodes per face in the face incidence: the facet's vertices (HEX20:
        # its 4 corners and 4 edge nodes)
        fnodes = 8 if self.ele_type == "HEX20" else len(element_info(self.ele_type)["faces"][0])
        finc = np.zeros(max(fnodes * len(allf), 1), dtype=np.int32)
        nat.check(getattr(L, f"fx_{tag}_face_incidence")(N, E, nat.ptr(self.cells), len(allf),
                                                         nat.ptr(self.all_faces), nat.ptr(finc_ptr), nat.ptr(finc)),
                  f"fx_{tag}_face_incidence")
        self.face_inc_ptr, self.face_inc = finc_ptr, finc
        max_deg = 96 if self.ele_type == "HEX20" else 64
        if self.csr_max_degree > max_deg:
            raise NotImplementedError(f"node degree {self.csr_max_degree} > {max_deg} not supported by the "
                                      "assembly kernel")

    @property
    def csr_row_of_slot(self):
        """Row of each CSR slot (nse,) int32 (feax/problem.py:436)."""
        if getattr(self, "_row_of_slot", None) is None:
            self._row_of_slot = np.repeat(np.arange(self.num_total_dofs_all_vars, dtype=np.int32),
                                          np.diff(self.csr_indptr))
        return self._row_of_slot

    @property
    def csr_diag_slots(self):
        """Slot of (d, d) for every row, -1 if absent (feax/problem.py:372-378)."""
        if getattr(self, "_diag_slots", None) is None:
            rows = self.csr_row_of_slot
            pos = np.nonzero(rows == self.csr_indices)[0]
            d = np.full(self.num_total_dofs_all_vars, -1, dtype=np.int32)
            d[rows[pos]] = pos
            self._diag_slots = d
        return self._diag_slots

    @property
    def csr_T_perm(self):
        """(nse,) int32: slot order of the transpose, A^T.data = data[csr_T_perm]
        (feax/problem.py:417-447, stable sort by (col, row))."""
        if getattr(self, "_T", None) is None:
            self._build_transpose()
        return self._T[0]

    @property
    def csr_T_indptr(self):
        if getattr(self, "_T", None) is None:
            self._build_transpose()
        return self._T[1]

    @property
    def csr_T_indices(self):
        if getattr(self, "_T", None) is None:
            self._build_transpose()
        return self._T[2]

    def _build_transpose(self):
        n = self.num_total_dofs_all_vars
        rows = self.csr_row_of_slot.astype(np.int64)
        cols = self.csr_indices.astype(np.int64)
        order = np.argsort(cols * n + rows, kind="stable")
        t_rows = cols[order]
        indptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(indptr, t_rows + 1, 1)
        self._T = (order.astype(np.int32), np.cumsum(indptr).astype(np.int32), rows[order].astype(np.int32))

    @property
    def csr_volume_slots(self):
        """(E, (vec npe)^2) int32: CSR slot of each element Jacobian entry
        jac[a*vec+i, b*vec+j] (feax/problem.py:458-463), from the rank table:
        indptr[vec n_a + i] + vec rank[e, a, b] + j."""
        if getattr(self, "_vol_slots", None) is None and self.coupled:
            # row DOF (a, i) / (a, T), column (b, j) / (b, T): u columns at
            # 3 rank + j, T columns after the 3 deg u columns
            c = self.cells.astype(np.int64)
            inds = self._element_inds(self.cells).astype(np.int64)                       # (E, 32)
            rank = self.elem_rank.reshape(-1, 8, 8).astype(np.int64)
            deg = (np.diff(self.csr_indptr)[3 * c] // 4).astype(np.int64)                 # (E, a)
            rowp = self.csr_indptr.astype(np.int64)[inds]                                 # (E, 32)
            node_of = np.concatenate([np.repeat(np.arange(8), 3), np.arange(8)])          # local node of a DOF
            comp_of = np.concatenate([np.tile(np.arange(3), 8), np.full(8, 3)])
            ra = rank[:, node_of[:, None], node_of[None, :]]                              # (E, 32, 32)
            colpos = np.where(comp_of[None, None, :] < 3, 3 * ra + comp_of[None, None, :],
                              3 * deg[:, node_of][:, :, None] + ra)
            self._vol_slots = (rowp[:, :, None] + colpos).reshape(len(c), 1024).astype(np.int32)
        if getattr(self, "_vol_slots", None) is None:
            c = self.cells.astype(np.int64)
            v = self.vec
            rank = self.elem_rank.reshape(-1, self.npe, self.npe).astype(np.int64)
            rowp = self.csr_indptr.astype(np.int64)[v * c[:, :, None] + np.arange(v)]      # (E, a, i)
            slot = (rowp[:, :, :, None, None] + v * rank[:, :, None, :, None]
                    + np.arange(v)[None, None, None, None, :])                            # (E, a, i, b, j)
            self._vol_slots = slot.reshape(len(c), (v * self.npe) ** 2).astype(np.int32)
        return self._vol_slots

    def _element_inds(self, cells):
        """Per-element DOFs, variable by variable (feax/problem.py:161-181)."""
        c = np.asarray(cells, np.int64)
        parts = [(fe.vec * c[:, :, None] + np.arange(fe.vec)[None, None, :] + off).reshape(len(c), -1)
                 for fe, off in zip(self.fes, self.offset)]
        return np.concatenate(parts, axis=1).astype(np.int32)

    def _coo(self):
        inds = [self._element_inds(self.cells)] + [self._element_inds(self.cells[b[:, 0]])
                                                   for b in self.boundary_inds_list]
        nd = inds[0].shape[1]
        I = np.concatenate([np.repeat(i[:, :, None], nd, axis=2).reshape(-1) for i in inds])
        J = np.concatenate([np.repeat(i[:, None, :], nd, axis=1).reshape(-1) for i in inds])
        return I, J

    @property
    def I(self):
        """Raw COO rows, volume then faces (feax/problem.py:183-212); built on demand."""
        if getattr(self, "_I", None) is None:
            self._I, self._J = self._coo()
        return self._I

    @property
    def J(self):
        if getattr(self, "_J", None) is None:
            self._I, self._J = self._coo()
        return self._J

    def _filter_mask(self):
        if self.matrix_view == MatrixView.UPPER:
            return self.J >= self.I
        if self.matrix_view == MatrixView.LOWER:
            return self.J <= self.I
        return None

    @property
    def I_filtered(self):
        m = self._filter_mask()
        return self.I if m is None else self.I[m]

    @property
    def J_filtered(self):
        m = self._filter_mask()
        return self.J if m is None else self.J[m]

    @property
    def filter_indices(self):
        m = self._filter_mask()
        return None if m is None else np.where(m)[0].astype(np.int32)

    def view_pattern(self):
        """(indptr, indices, full_slot_of_view_slot) of the matrix_view
        pattern: the FULL pattern filtered to j >= i (UPPER) or j <= i
        (LOWER), which is what feax's dedupe of the filtered (I, J) gives."""
        if self.matrix_view == MatrixView.FULL:
            return self.csr_indptr, self.csr_indices, None
        rows = self.csr_row_of_slot
        keep = self.csr_indices >= rows if self.matrix_view == MatrixView.UPPER else self.csr_indices <= rows
        slots = np.nonzero(keep)[0].astype(np.int32)
        counts = np.bincount(rows[slots], minlength=self.num_total_dofs_all_vars)
        indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        return indptr, self.csr_indices[slots], slots

    # ------------------------------------------------------------------ layout
    def unflatten_fn_sol_list(self, flat):
        flat = getattr(flat, "dofs", flat)
        out, start = [], 0
        for nn, v in self.solution_layout:
            out.append(flat[start:start + nn * v].reshape(nn, v))
            start += nn * v
        return out

    # ------------------------------------------------------------------ hooks
    def custom_init(self, *args):
        pass

    def get_tensor_map(self):
        return None

    def get_energy_density(self):
        return None

    def get_surface_maps(self):
        return []

    def get_mass_map(self):
        return None

    def get_weak_form(self):
        return None
