"""Oracle: TET4 geometry, element kernels and global assembly (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:
  feax/basis.py:78-83           TET4: basix P1 tetrahedron, vertex order
                                kept (re_order [0, 1, 2, 3]), gauss_order 0:
                                one point at the centroid, weight 1/6
  feax/basis.py:244-325         facets: basix's tetrahedron facets (facet i
                                opposite vertex i), one point at the facet
                                centroid, weight 1/2 * |facet Jacobian|;
                                outward reference normals
  feax/fe.py:96-129, 175-239    J = sum_a x_a (x) grad_xi N_a, grad N =
                                grad_xi N J^-1, JxW = det J * w (no abs);
                                Nanson scale ||n_ref J^-1|| det J w_face
  feax/assembler.py:406-448, 494-535, 808-818
                                element residual / Jacobian / dead traction,
                                as for HEX8 with one quadrature point
  feax/fe.py:267-318            boundary faces: all facet vertices satisfy
                                location_fn, (cell, facet) argwhere order

Pinned by (tests/test_oracle_tet4.py): rigid-body modes in the kernel of
every element matrix; the constant-strain patch test (interior residual
zero, boundary reactions = sigma n area); the strain energy 1/2 u^T K u =
1/2 V sigma:eps of a linear field; the total dead traction t * area; the
finite-difference tangent of the Neo-Hookean residual; a scipy direct
solve of the assembled system.
"""

import numpy as np

from . import hex8
from . import structure as st

TET4_FACE_INDS = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]], dtype=np.int64)
REF_GRADS = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
# basix reference facet data: outward normals and |facet Jacobian|
FACE_NORMALS = np.array([[1.0, 1.0, 1.0], [-1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, -1.0]])
FACE_NORMALS[0] /= np.sqrt(3.0)
FACE_JAC = np.array([np.sqrt(3.0), 1.0, 1.0, 1.0])


def box_tets(dims, spacing=(1.0, 1.0, 1.0)):
    """A conforming TET4 mesh of the StructuredGrid box: every hexahedral cell
    split into the 6 tetrahedra of its main diagonal (one per ordering of the
    three axes), vertices ordered so that det J > 0.  Returns (points, cells)."""
    nx, ny, nz = dims
    pts, _ = st.structured_grid(dims, spacing)

    def node(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    cells = []
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for p in perms:
                    v = [0, 0, 0]
                    path = [node(i, j, k)]
                    for ax in p:
                        v[ax] += 1
                        path.append(node(i + v[0], j + v[1], k + v[2]))
                    cells.append(path)
    cells = np.asarray(cells, np.int64)
    X = pts[cells]
    det = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
    neg = det < 0
    cells[neg, 2], cells[neg, 3] = cells[neg, 3].copy(), cells[neg, 2].copy()
    return pts, cells


def boundary_faces(points, cells, location_fn):
    """(F, 2) (cell, facet) whose 3 vertices satisfy location_fn."""
    fp = points[cells[:, TET4_FACE_INDS]]                  # (E, 4, 3, 3)
    on = np.vectorize(lambda *x: bool(location_fn(np.array(x))), signature="(),(),()->()")(
        fp[..., 0], fp[..., 1], fp[..., 2])
    return np.argwhere(on.all(axis=2)).astype(np.int64)


def geometry(points, cells):
    """grads (E, 1, 4, 3) and JxW (E, 1)."""
    X = points[cells]                                       # (E, 4, 3)
    J = np.einsum("eai,aj->eij", X, REF_GRADS)
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)
    grads = np.einsum("aj,ejd->ead", REF_GRADS, Ji)
    return grads[:, None], (det / 6.0)[:, None]


def face_geometry(points, cells, boundary_inds):
    """face shape values (F, 1, 4) at the facet centroid and Nanson scale (F, 1)."""
    e, lf = boundary_inds[:, 0], boundary_inds[:, 1]
    X = points[cells[e]]
    J = np.einsum("fai,aj->fij", X, REF_GRADS)
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)
    m = np.einsum("fj,fjd->fd", FACE_NORMALS[lf], Ji)
    nanson = np.linalg.norm(m, axis=1) * det * 0.5 * FACE_JAC[lf]
    vals = np.zeros((len(e), 4))
    for a in range(4):
        vals[:, a] = (TET4_FACE_INDS[lf] == a).any(axis=1) / 3.0
    return vals[:, None, :], nanson[:, None]


def param_at_quads(param, cells, num_nodes):
    """(E, 1): node var -> N(centroid) . v_e = mean of the 4 values; cell var as is."""
    param = np.asarray(param, float)
    if param.ndim == 0:
        return np.full((len(cells), 1), float(param))
    if param.shape[0] == num_nodes:
        return (param[cells] * 0.25).sum(axis=1)[:, None]
    if param.shape[0] == len(cells):
        return param[:, None]
    raise ValueError(f"unsupported volume var shape {param.shape}")


def element_residual(u_e, grads, JxW, pq, mat):
    ug = np.einsum("eai,eqad->eqid", u_e, grads)
    sig = mat.stress(ug, pq)
    return np.einsum("eqid,eqad,eq->eai", sig, grads, JxW).reshape(len(u_e), 12)


def element_stiffness(u_e, grads, JxW, pq, mat):
    n = len(u_e)
    if mat.kind == "neo_hookean":
        ug = np.einsum("eai,eqad->eqid", u_e, grads)
        C = mat.tangent(ug, pq)
        return np.einsum("eqad,eqidjl,eqbl,eq->eaibj", grads, C, grads, JxW).reshape(n, 12, 12)
    lam, mu = hex8.lame(mat.modulus(pq), mat.nu)
    wl, wm = lam * JxW, mu * JxW
    K = (np.einsum("eqai,eqbj,eq->eaibj", grads, grads, wl)
         + np.einsum("eqaj,eqbi,eq->eaibj", grads, grads, wm))
    gg = np.einsum("eqad,eqbd,eq->eab", grads, grads, wm)
    K += gg[:, :, None, :, None] * np.eye(3)[None, None, :, None, :]
    return K.reshape(n, 12, 12)


def constant_traction(boundary_inds, magnitude, direction=(0.0, 0.0, 1.0)):
    """(F, 1, 3) dead traction direction * magnitude."""
    m = np.broadcast_to(np.asarray(magnitude, float), (len(boundary_inds), 1))
    return m[:, :, None] * np.asarray(direction, float)[None, None, :]


class Model:
    """The oracle for one TET4 vec = 3 problem (the interface of hex8.Model)."""

    def __init__(self, points, cells, boundary_inds_list=(), view="FULL"):
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.E = len(self.cells)
        self.N = len(self.points)
        self.n = 3 * self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = geometry(self.points, self.cells)
        self.faces = [face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells)
        I, J, self.res_dofs = st.coo_ij(self.cells, self.boundary_inds_list)
        If, Jf, self.filter_indices = st.filter_view(I, J, view)
        self.csr = st.csr_structure(If, Jf, self.n, I, J, num_cells=self.E, ndof_vol=12)
        self.nse = self.csr["nse"]

    def residual(self, u, param, mat, tractions=()):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        vals = [element_residual(ue, self.grads, self.JxW, pq, mat).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, t in zip(self.faces, self.boundary_inds_list, tractions):
            vals.append(np.einsum("fqi,fqa,fq->fai", t, fv, ns).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals), minlength=self.n)

    def element_matrices(self, u, param, mat):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        return element_stiffness(ue, self.grads, self.JxW, pq, mat)

    def csr_values(self, u, param, mat):
        Ke = self.element_matrices(u, param, mat)
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        return np.bincount(slots, Ke.reshape(-1), minlength=self.nse + 1)[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))
