"""Oracle: HEX20 basis, geometry, element kernels and assembly (FP64 numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates feax's HEX20 (feax/basis.py:71-77, 155-197, 244-325): basix's
serendipity element of degree 2 on the hexahedron, re-ordered to the
meshio / Abaqus node order by [0,1,3,2,4,5,7,6,8,11,13,9,16,18,19,17,10,12,
15,14]; quadrature basix.make_quadrature(hexahedron, 2) = the 2x2x2 Gauss
rule (weights 1/8 on [0,1]^3), faces 2x2 Gauss per facet (weights 1/4);
boundary faces are selected by their 4 corner vertices (the facet vertex
lists through the re-order).  Geometry, element residual / Jacobian,
surface traction, CSR accumulation and the Dirichlet elimination are the
HEX8 formulas (oracle/hex8.py) with 20 nodes (60 DOFs) per element.

The basis.  basix's serendipity element takes as degrees of freedom the
values at the 8 vertices and, on each of the 12 edges, the integral moment
against the degree-0 polynomial, i.e. the mean of the function over the
edge (the edges of the reference cell have length 1).  Its basis functions
are written here from the 20-node nodal serendipity functions psi (values
at vertices and edge midpoints) as
    phi_edge   = 3/2 psi_edge                        (mean 1 on its edge),
    phi_vertex = psi_vertex - 1/4 sum_{e at v} psi_e (mean 0 on every edge).
The edge DOF of a field is therefore its edge mean (for a field linear
along the edge, the midpoint value).  Parity note: basix is absent here;
that its default serendipity variant uses these moment DOFs is basix's
documented construction, taken as given (parity unpinned for the variant).

Pinned by (tests/test_oracle_hex20.py): the dual-basis identities (vertex
values, edge means by 3-point Gauss on each edge), partition of unity,
exact reproduction of every quadratic (DOFs -> field), reference
gradients against central differences, rigid-body modes in the kernel of
the element stiffness and symmetry, the linear patch test on a distorted
mesh, the strain energy of a homogeneous strain, and the consistent nodal
loads of a uniform traction on a flat face (corners -A/4, edges A/2 each).
"""

import numpy as np

from . import hex8

# feax HEX20 local nodes: meshio corners (hex8.XI_NODES) then the midpoints
# of the meshio edges 8..19.
EDGES = np.array([(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                  (0, 4), (1, 5), (2, 6), (3, 7)], dtype=np.int64)
XI_NODES = np.concatenate([hex8.XI_NODES, 0.5 * (hex8.XI_NODES[EDGES[:, 0]] + hex8.XI_NODES[EDGES[:, 1]])])
QUAD_POINTS = hex8.QUAD_POINTS
QUAD_WEIGHTS = hex8.QUAD_WEIGHTS
FACE_QUAD_POINTS = hex8.FACE_QUAD_POINTS
FACE_WEIGHTS = hex8.FACE_WEIGHTS
FACE_NORMALS = hex8.FACE_NORMALS
NPE = 20
NDOF = 60


def _edges_at_vertex():
    return [[8 + k for k, (a, b) in enumerate(EDGES) if v in (a, b)] for v in range(8)]


EDGES_AT = _edges_at_vertex()


def nodal_serendipity(xi):
    """psi (M, 20) and d psi / d xi (M, 20, 3) of the 20-node nodal
    serendipity hexahedron (values 1 at their own vertex / edge midpoint)."""
    xi = np.atleast_2d(np.asarray(xi, float))
    s = 2.0 * xi - 1.0                                   # (M, 3) in [-1, 1]
    sig = 2.0 * XI_NODES - 1.0                           # corners +-1, edge midpoints 0 on their axis
    M = len(xi)
    psi = np.zeros((M, 20))
    dpsi = np.zeros((M, 20, 3))
    for a in range(8):
        f = 1.0 + sig[a] * s                             # (M, 3)
        g = (sig[a] * s).sum(axis=1) - 2.0
        P = f.prod(axis=1)
        psi[:, a] = P * g / 8.0
        for k in range(3):
            others = np.prod(np.delete(f, k, axis=1), axis=1)
            dpsi[:, a, k] = (sig[a, k] * others * g + P * sig[a, k]) / 8.0 * 2.0
    for e in range(12):
        a = 8 + e
        k = int(np.flatnonzero(sig[a] == 0)[0])          # the edge's axis
        i, j = [d for d in range(3) if d != k]
        fi = 1.0 + sig[a, i] * s[:, i]
        fj = 1.0 + sig[a, j] * s[:, j]
        q = 1.0 - s[:, k] ** 2
        psi[:, a] = 0.25 * q * fi * fj
        dpsi[:, a, k] = 0.25 * (-2.0 * s[:, k]) * fi * fj * 2.0
        dpsi[:, a, i] = 0.25 * q * sig[a, i] * fj * 2.0
        dpsi[:, a, j] = 0.25 * q * fi * sig[a, j] * 2.0
    return psi, dpsi


def shape_vals_grads(xi):
    """phi (M, 20) and d phi / d xi (M, 20, 3): basix serendipity degree 2
    (vertex values, edge means), feax local order."""
    psi, dpsi = nodal_serendipity(xi)
    phi = psi.copy()
    dphi = dpsi.copy()
    phi[:, 8:] *= 1.5
    dphi[:, 8:] *= 1.5
    for v in range(8):
        for e in EDGES_AT[v]:
            phi[:, v] -= 0.25 * psi[:, e]
            dphi[:, v] -= 0.25 * dpsi[:, e]
    return phi, dphi


SHAPE_VALS, SHAPE_GRADS_REF = shape_vals_grads(QUAD_POINTS)
FACE_SHAPE_VALS, FACE_SHAPE_GRADS_REF = (
    a.reshape(6, 4, *a.shape[1:]) for a in shape_vals_grads(FACE_QUAD_POINTS.reshape(-1, 3)))
# Nodes of each facet (4 corners then the 4 edges between them), and the
# corner list used to select boundary faces (feax/basis.py:321-324).
FACE_CORNERS = np.array([[0, 1, 3, 2], [0, 1, 4, 5], [0, 3, 4, 7], [1, 2, 5, 6], [3, 2, 7, 6], [4, 5, 7, 6]],
                        dtype=np.int64)
FACE_NODES = np.array([list(fc) + [8 + k for k, (a, b) in enumerate(EDGES) if a in fc and b in fc]
                       for fc in FACE_CORNERS], dtype=np.int64)


def geometry(points, cells):
    """Physical shape gradients (E, 8q, 20a, 3) and JxW (E, 8q) -- fe.py:96-129."""
    X = np.asarray(points, float)[np.asarray(cells)]
    Jm = np.einsum("eai,qaj->eqij", X, SHAPE_GRADS_REF)
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)
    grads = np.einsum("qaj,eqji->eqai", SHAPE_GRADS_REF, inv)
    return grads, det * QUAD_WEIGHTS[None, :]


def face_geometry(points, cells, boundary_inds):
    """Face shape values (F, 4, 20) and nanson scale (F, 4) -- fe.py:175-222."""
    X = np.asarray(points, float)[np.asarray(cells)[boundary_inds[:, 0]]]
    fidx = boundary_inds[:, 1]
    g = FACE_SHAPE_GRADS_REF[fidx]
    Jm = np.einsum("fai,fqaj->fqij", X, g)
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)
    n = FACE_NORMALS[fidx]
    nanson = np.linalg.norm(np.einsum("fj,fqji->fqi", n, inv), axis=-1)
    return FACE_SHAPE_VALS[fidx], nanson * det * FACE_WEIGHTS[fidx]


def param_at_quads(param, cells, num_nodes):
    """node var -> phi . v_e at the 8 Gauss points, cell var -> broadcast."""
    param = np.asarray(param, float)
    if param.ndim == 0:
        return np.full((len(cells), 8), float(param))
    if param.shape[0] == num_nodes:
        return np.einsum("qa,ea->eq", SHAPE_VALS, param[cells])
    if param.shape[0] == len(cells):
        return np.repeat(param[:, None], 8, axis=1)
    raise ValueError(f"unsupported volume var shape {param.shape}")


def element_residual(u_e, grads, JxW, pq, mat):
    """(E, 60) volume residual r_(a,i) = sum_q sigma_id dphi_a/dx_d JxW."""
    ug = np.einsum("eai,eqad->eqid", u_e, grads)
    sig = mat.stress(ug, pq)
    return np.einsum("eqid,eqad,eq->eai", sig, grads, JxW).reshape(len(u_e), NDOF)


def element_stiffness(u_e, grads, JxW, pq, mat):
    """(E, 60, 60) element Jacobian jac[a*3+i, b*3+j]."""
    n = len(u_e)
    if mat.kind == "neo_hookean":
        ug = np.einsum("eai,eqad->eqid", u_e, grads)
        C = mat.tangent(ug, pq)
        return np.einsum("eqad,eqidjl,eqbl,eq->eaibj", grads, C, grads, JxW).reshape(n, NDOF, NDOF)
    lam, mu = hex8.lame(mat.modulus(pq), mat.nu)
    wl, wm = lam * JxW, mu * JxW
    K = (np.einsum("eqai,eqbj,eq->eaibj", grads, grads, wl)
         + np.einsum("eqaj,eqbi,eq->eaibj", grads, grads, wm))
    gg = np.einsum("eqad,eqbd,eq->eab", grads, grads, wm)
    K += gg[:, :, None, :, None] * np.eye(3)[None, None, :, None, :]
    return K.reshape(n, NDOF, NDOF)


def face_residual(face_vals, nanson, traction):
    """(F, 60) surface residual for a dead traction t (F, 4, 3), + sign."""
    return np.einsum("fqi,fqa,fq->fai", traction, face_vals, nanson).reshape(len(face_vals), NDOF)


class Model(hex8.Model):
    """hex8.Model with 20 nodes / 60 DOFs per element."""

    def __init__(self, points, cells, boundary_inds_list=(), view="FULL", structure=True):
        from . import structure as st
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.E = len(self.cells)
        self.N = len(self.points)
        self.n = 3 * self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = geometry(self.points, self.cells)
        self.faces = [face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        self.inds = st.build_inds(self.cells)
        if structure:
            I, J, self.res_dofs = st.coo_ij(self.cells, self.boundary_inds_list)
            If, Jf, self.filter_indices = st.filter_view(I, J, view)
            self.csr = st.csr_structure(If, Jf, self.n, I, J, num_cells=self.E, ndof_vol=NDOF)
            self.nse = self.csr["nse"]
            self.vol_len = self.E * NDOF * NDOF

    def residual(self, u, param, mat, tractions=()):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        vals = [element_residual(ue, self.grads, self.JxW, pq, mat).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, t in zip(self.faces, self.boundary_inds_list, tractions):
            vals.append(face_residual(fv, ns, t).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals), minlength=self.n)

    def element_matrices(self, u, param, mat, chunk=2048):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        out = np.empty((self.E, NDOF, NDOF))
        for s in range(0, self.E, chunk):
            e = slice(s, s + chunk)
            out[e] = element_stiffness(ue[e], self.grads[e], self.JxW[e], pq[e], mat)
        return out


def box(dims, spacing=(1.0, 1.0, 1.0)):
    """(points (N, 3), cells (E, 20) int64) of a structured HEX20 box: the
    corners of each cell in the corner order of the HEX8 StructuredGrid
    (feax/spgrid.py:21-25) and one node at the midpoint of every edge,
    numbered after the corners in order of first appearance."""
    from . import structure as st
    pts8, cells8 = st.structured_grid(dims, spacing)
    pairs = cells8[:, EDGES]                                  # (E, 12, 2)
    key = np.sort(pairs, axis=2)
    flat = key.reshape(-1, 2)
    uniq, first, inv = np.unique(flat[:, 0] * len(pts8) + flat[:, 1], return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")                  # first appearance
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    mid_ids = len(pts8) + rank[inv].reshape(len(cells8), 12)
    e0 = flat[first[order], 0]
    e1 = flat[first[order], 1]
    mids = 0.5 * (pts8[e0] + pts8[e1])
    return np.concatenate([pts8, mids]), np.concatenate([cells8, mid_ids], axis=1)


def boundary_faces(points, cells, location_fn):
    """(F, 2) (cell, face) whose 4 corner vertices satisfy location_fn,
    evaluated per node (hex8 corner lists)."""
    from .structure import _location_per_node
    flag = _location_per_node(points, location_fn)
    return np.argwhere(flag[cells[:, FACE_CORNERS]].all(axis=2)).astype(np.int64)
