"""Oracle: scalar HEX8 conduction (vec = 1), FP64 numpy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, for feax's scalar Problem with the flux tensor map
``lambda grad_T, k: k * grad_T`` (tests/test_params_bridge.py:131-133,
examples/basic/thermo_mech_coupling.py:64-69):
  feax/assembler.py:406-448     R_a = sum_q k_q grad N_a . grad T_q JxW_q
  feax/assembler.py:808-818     K_ab = sum_q k_q grad N_a . grad N_b JxW_q
  feax/assembler.py:494-535     surface map q (one component): + sum_q q N_a nanson
  feax/problem.py:161-212       DOF a = node a (vec = 1), COO order
with the HEX8 geometry, Gauss rules and boundary faces of oracle/hex8.py.

Pinned by (tests/test_oracle_conduction.py): constants span the kernel of
every element matrix (rank 7); the linear-field patch test on a distorted
mesh (interior residual zero, face reactions k g.n area); the energy
1/2 T^T K T = 1/2 V k |g|^2; the total prescribed flux q * area; the exact
linear profile of a bar between two Dirichlet faces.
"""

import numpy as np

from . import hex8
from . import structure as st


def conductivity(kind, p, k0=1.0, kmin=1e-3, penal=3.0):
    if kind == "simp":
        return kmin + np.asarray(p) ** penal * (k0 - kmin)
    return np.asarray(p, float)


class Model:
    """The oracle for one scalar HEX8 problem; kind "conduction" (k = p) or
    "simp" (k = kmin + p^penal (k0 - kmin))."""

    def __init__(self, points, cells, boundary_inds_list=(), view="FULL"):
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.E, self.N = len(self.cells), len(self.points)
        self.n = self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = hex8.geometry(self.points, self.cells)          # (E, 8, 8, 3), (E, 8)
        self.faces = [hex8.face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells, vec=1)
        I, J, self.res_dofs = st.coo_ij(self.cells, self.boundary_inds_list, vec=1)
        If, Jf, _ = st.filter_view(I, J, view)
        self.csr = st.csr_structure(If, Jf, self.n, I, J, num_cells=self.E, ndof_vol=8)
        self.nse = self.csr["nse"]

    def _k(self, param, kind, **law):
        return conductivity(kind, hex8.param_at_quads(param, self.cells, self.N), **law)   # (E, 8)

    def element_matrices(self, param, kind="conduction", **law):
        kw = self._k(param, kind, **law) * self.JxW
        return np.einsum("eqad,eqbd,eq->eab", self.grads, self.grads, kw)

    def residual(self, T, param, kind="conduction", fluxes=(), **law):
        kw = self._k(param, kind, **law) * self.JxW
        gT = np.einsum("ea,eqad->eqd", T[self.cells], self.grads)
        vals = [np.einsum("eqd,eqad,eq->ea", gT, self.grads, kw).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, q in zip(self.faces, self.boundary_inds_list, fluxes):
            vals.append(np.einsum("fq,fqa,fq->fa", np.broadcast_to(q, ns.shape), fv, ns).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals), minlength=self.n)

    def csr_values(self, param, kind="conduction", **law):
        Ke = self.element_matrices(param, kind, **law)
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        return np.bincount(slots, Ke.reshape(-1), minlength=self.nse + 1)[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))


class QuadModel:
    """The scalar oracle on QUAD4 (plane conduction): oracle/quad4.py's
    geometry (2 x 2 Gauss points, 2 per edge) with the flux law above."""

    def __init__(self, points, cells, boundary_inds_list=()):
        from . import quad4
        self.points = np.asarray(points, float)[:, :2]
        self.cells = np.asarray(cells, np.int64)
        self.E, self.N = len(self.cells), len(self.points)
        self.n = self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = quad4.geometry(self.points, self.cells)            # (E, 4, 4, 2), (E, 4)
        self.faces = [quad4.face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells, vec=1)
        I, J, _ = st.coo_ij(self.cells, self.boundary_inds_list, vec=1)
        self.csr = st.csr_structure(I, J, self.n, I, J, num_cells=self.E, ndof_vol=4)
        self.nse = self.csr["nse"]
        self._pq = quad4.param_at_quads

    def _kw(self, param, kind, **law):
        return conductivity(kind, self._pq(param, self.cells, self.N), **law) * self.JxW

    def element_matrices(self, param, kind="conduction", **law):
        return np.einsum("eqad,eqbd,eq->eab", self.grads, self.grads, self._kw(param, kind, **law))

    def residual(self, T, param, kind="conduction", fluxes=(), **law):
        gT = np.einsum("ea,eqad->eqd", T[self.cells], self.grads)
        vals = [np.einsum("eqd,eqad,eq->ea", gT, self.grads, self._kw(param, kind, **law)).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, q in zip(self.faces, self.boundary_inds_list, fluxes):
            vals.append(np.einsum("fq,fqa,fq->fa", np.broadcast_to(q, ns.shape), fv, ns).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals), minlength=self.n)

    def csr_values(self, param, kind="conduction", **law):
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        return np.bincount(slots, self.element_matrices(param, kind, **law).reshape(-1), minlength=self.nse + 1)[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))


class TetModel(QuadModel):
    """The scalar oracle on TET4: oracle/tet4.py's geometry (one Gauss point,
    facet flux at the facet centroid) with the flux law above."""

    def __init__(self, points, cells, boundary_inds_list=()):
        from . import tet4
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.E, self.N = len(self.cells), len(self.points)
        self.n = self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = tet4.geometry(self.points, self.cells)            # (E, 1, 4, 3), (E, 1)
        self.faces = [tet4.face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells, vec=1)
        I, J, _ = st.coo_ij(self.cells, self.boundary_inds_list, vec=1)
        self.csr = st.csr_structure(I, J, self.n, I, J, num_cells=self.E, ndof_vol=4)
        self.nse = self.csr["nse"]
        self._pq = tet4.param_at_quads
