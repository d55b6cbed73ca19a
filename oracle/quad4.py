"""Oracle: QUAD4 (dim 2, vec 2) geometry, element kernels and global assembly
(FP64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:
  feax/basis.py:90-95           QUAD4: basix P1 quadrilateral re-ordered
                                [0, 1, 3, 2] -> local nodes (0,0), (1,0),
                                (1,1), (0,1); gauss_order 2: 2 x 2 Gauss
                                points, first coordinate slowest, weight 1/4
  feax/basis.py:244-325         edges: basix's quadrilateral facets under the
                                re-order -> [0,1], [0,3], [1,2], [3,2], each
                                with 2 Gauss points (weight 1/2 x edge length
                                1) mapped v0 + s (v1 - v0), outward
                                reference normals (0,-1), (-1,0), (1,0), (0,1)
  feax/fe.py:96-129, 175-239    J = sum_a x_a (x) grad_xi N_a, grad N =
                                grad_xi N J^-1, JxW = det J * w (no abs);
                                Nanson scale ||n_ref J^-1|| det J w_face
  feax/assembler.py:406-448, 494-535, 808-818
                                element residual / Jacobian / dead traction
                                with 2 x 2 tensors
  feax/mesh.py:447-502          rectangle_mesh node and cell numbering

The material laws are written with 2 x 2 tensors (plane strain:
sigma = lambda tr(eps) I + 2 mu eps; Neo-Hookean P = mu (F - F^-T) +
kappa J (J - 1) F^-T with F = I + grad u, J = det F), independently of the
kernels' 3-D embedding.

Pinned by (tests/test_oracle_quad4.py): the three rigid-body modes span
the kernel of every element matrix; the constant-strain patch test on a
distorted mesh (interior residual zero, edge reactions sigma n length); the
strain energy 1/2 u^T K u = 1/2 A sigma:eps of a linear field; the total
dead traction t * length and the half-length share of each edge vertex; the
finite-difference tangent of the Neo-Hookean residual; the Neo-Hookean
tangent at F = I equal to the linear-elastic one with (lambda, mu) =
(kappa, mu).
"""

import numpy as np

from . import hex8
from . import structure as st

QUAD4_FACE_INDS = np.array([[0, 1], [0, 3], [1, 2], [3, 2]], dtype=np.int64)
REF_NODES = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
FACE_NORMALS = np.array([[0.0, -1.0], [-1.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
_G = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
QUAD_POINTS = np.array([[_G[i], _G[j]] for i in range(2) for j in range(2)])


def shape(xi):
    """N (P, 4) and reference gradients (P, 4, 2) of the bilinear basis."""
    xi = np.atleast_2d(xi)
    s, t = xi[:, 0:1], xi[:, 1:2]
    fs = np.where(REF_NODES[:, 0] > 0.5, s, 1.0 - s)
    ft = np.where(REF_NODES[:, 1] > 0.5, t, 1.0 - t)
    ss = np.where(REF_NODES[:, 0] > 0.5, 1.0, -1.0)
    st_ = np.where(REF_NODES[:, 1] > 0.5, 1.0, -1.0)
    return fs * ft, np.stack([ss * ft, fs * st_], axis=2)


SHAPE_VALS, REF_GRADS = shape(QUAD_POINTS)          # (4, 4), (4, 4, 2)


def rectangle(nx, ny, lx=1.0, ly=1.0):
    """feax's rectangle_mesh numbering: node i (ny + 1) + j at (x_i, y_j)."""
    x, y = np.linspace(0.0, lx, nx + 1), np.linspace(0.0, ly, ny + 1)
    pts = np.array([[xi, yj] for xi in x for yj in y])
    cells = []
    for i in range(nx):
        for j in range(ny):
            n0 = i * (ny + 1) + j
            cells.append([n0, n0 + ny + 1, n0 + ny + 2, n0 + 1])
    return pts, np.asarray(cells, np.int64)


def boundary_faces(points, cells, location_fn):
    """(F, 2) (cell, edge) whose 2 vertices satisfy location_fn."""
    ep = points[cells[:, QUAD4_FACE_INDS]]                  # (E, 4, 2, 2)
    on = np.vectorize(lambda *x: bool(location_fn(np.array(x))), signature="(),()->()")(ep[..., 0], ep[..., 1])
    return np.argwhere(on.all(axis=2)).astype(np.int64)


def _jac(points, cells, ref_grads):
    X = points[cells]                                       # (E, 4, 2)
    return np.einsum("eai,qaj->eqij", X, ref_grads)


def geometry(points, cells):
    """grads (E, 4, 4, 2) and JxW (E, 4)."""
    J = _jac(points, cells, REF_GRADS)
    Ji = np.linalg.inv(J)
    grads = np.einsum("qaj,eqjd->eqad", REF_GRADS, Ji)
    return grads, np.linalg.det(J) * 0.25


def face_points(lf):
    v0, v1 = REF_NODES[QUAD4_FACE_INDS[lf, 0]], REF_NODES[QUAD4_FACE_INDS[lf, 1]]
    return v0[None] + _G[:, None] * (v1 - v0)[None]         # (2, 2)


def face_geometry(points, cells, boundary_inds):
    """Edge shape values (F, 2, 4) and Nanson scale (F, 2)."""
    F = len(boundary_inds)
    vals, nanson = np.zeros((F, 2, 4)), np.zeros((F, 2))
    for k, (e, lf) in enumerate(boundary_inds):
        N, dN = shape(face_points(lf))
        J = np.einsum("ai,qaj->qij", points[cells[e]], dN)
        m = np.einsum("j,qjd->qd", FACE_NORMALS[lf], np.linalg.inv(J))
        vals[k] = N
        nanson[k] = np.linalg.norm(m, axis=1) * np.linalg.det(J) * 0.5
    return vals, nanson


def param_at_quads(param, cells, num_nodes):
    """(E, 4): node var -> N . v_e, cell var -> broadcast."""
    param = np.asarray(param, float)
    if param.ndim == 0:
        return np.full((len(cells), 4), float(param))
    if param.shape[0] == num_nodes:
        return np.einsum("qa,ea->eq", SHAPE_VALS, param[cells])
    if param.shape[0] == len(cells):
        return np.repeat(param[:, None], 4, axis=1)
    raise ValueError(f"unsupported volume var shape {param.shape}")


def lin_stress(ug, E, nu):
    lam, mu = hex8.lame(E, nu)
    eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
    tr = np.trace(eps, axis1=-2, axis2=-1)
    return lam[..., None, None] * tr[..., None, None] * np.eye(2) + 2 * mu[..., None, None] * eps


def nh_stress(ug, mu, kappa):
    F = ug + np.eye(2)
    J = np.linalg.det(F)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    return mu[..., None, None] * (F - FiT) + (kappa * J * (J - 1.0))[..., None, None] * FiT


def nh_tangent(ug, mu, kappa):
    """dP_iJ/dF_kL (..., 2, 2, 2, 2)."""
    F = ug + np.eye(2)
    J = np.linalg.det(F)
    A = np.swapaxes(np.linalg.inv(F), -1, -2)
    I = np.eye(2)
    c1 = kappa * (2 * J - 1) * J
    c2 = mu - kappa * J * (J - 1)
    return (mu[..., None, None, None, None] * np.einsum("ik,JL->iJkL", I, I)
            + c1[..., None, None, None, None] * np.einsum("...iJ,...kL->...iJkL", A, A)
            + c2[..., None, None, None, None] * np.einsum("...iL,...kJ->...iJkL", A, A))


class Model:
    """The oracle for one QUAD4 vec = 2 problem (the interface of hex8.Model)."""

    def __init__(self, points, cells, boundary_inds_list=(), view="FULL"):
        self.points = np.asarray(points, float)[:, :2]
        self.cells = np.asarray(cells, np.int64)
        self.E = len(self.cells)
        self.N = len(self.points)
        self.n = 2 * self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = geometry(self.points, self.cells)
        self.faces = [face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells, vec=2)
        I, J, self.res_dofs = st.coo_ij(self.cells, self.boundary_inds_list, vec=2)
        If, Jf, self.filter_indices = st.filter_view(I, J, view)
        self.csr = st.csr_structure(If, Jf, self.n, I, J, num_cells=self.E, ndof_vol=8)
        self.nse = self.csr["nse"]

    def _ug(self, u):
        return np.einsum("eai,eqad->eqid", u.reshape(-1, 2)[self.cells], self.grads)

    def _stress(self, ug, pq, mat):
        if mat.kind == "neo_hookean":
            return nh_stress(ug, pq, mat.kappa)
        return lin_stress(ug, mat.modulus(pq), mat.nu)

    def residual(self, u, param, mat, tractions=()):
        pq = param_at_quads(param, self.cells, self.N)
        sig = self._stress(self._ug(u), pq, mat)
        vals = [np.einsum("eqid,eqad,eq->eai", sig, self.grads, self.JxW).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, t in zip(self.faces, self.boundary_inds_list, tractions):
            vals.append(np.einsum("fqi,fqa,fq->fai", t, fv, ns).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals), minlength=self.n)

    def element_matrices(self, u, param, mat):
        pq = param_at_quads(param, self.cells, self.N)
        if mat.kind == "neo_hookean":
            C = nh_tangent(self._ug(u), pq, mat.kappa)
        else:
            lam, mu = hex8.lame(mat.modulus(pq), mat.nu)
            I = np.eye(2)
            C = (lam[..., None, None, None, None] * np.einsum("ij,kl->ijkl", I, I)
                 + mu[..., None, None, None, None] * (np.einsum("ik,jl->ijkl", I, I)
                                                      + np.einsum("il,jk->ijkl", I, I)))
        K = np.einsum("eqad,eqidjl,eqbl,eq->eaibj", self.grads, C, self.grads, self.JxW)
        return K.reshape(self.E, 8, 8)

    def csr_values(self, u, param, mat):
        Ke = self.element_matrices(u, param, mat)
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        return np.bincount(slots, Ke.reshape(-1), minlength=self.nse + 1)[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))


def constant_traction(boundary_inds, magnitude, direction=(0.0, 1.0)):
    """(F, 2, 2) dead traction direction * magnitude."""
    m = np.broadcast_to(np.asarray(magnitude, float), (len(boundary_inds), 2))
    return m[:, :, None] * np.asarray(direction, float)[None, None, :]
