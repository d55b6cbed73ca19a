"""Oracle: HEX8 basis, geometry, element kernels and assembly (FP64 numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:
  feax/basis.py:155-197, 244-325   basix P1 hexahedron at 2x2x2 Gauss points
                                   (weights 1/8 on [0,1]^3), face quadrature
                                   2x2 per facet (weights 1/4), outward normals
  feax/fe.py:88, 96-129, 175-222   J = sum_a x_a (x) dN_a/dxi, dN/dx = dN/dxi J^-1,
                                   JxW = det(J) w, nanson = |n_ref J^-1| det(J) w
  feax/assembler.py:121-356        Operator.grad / interpolate_var / integrate_*
  feax/assembler.py:406-448        laplace kernel r_(a,i) = sum_q sigma_id dN_a/dx_d JxW
  feax/assembler.py:494-535        surface kernel r_a += sum_q t N_a nanson (+ sign)
  feax/assembler.py:808-818        element Jacobian jac[r, c] = d r_r / d u_c
  feax/assembler.py:1058-1103      residual = one ordered reduction over DOFs
  feax/assembler.py:1253-1318      CSR values = ordered reduction over slots
  feax/DCboundary.py:147-251       Dirichlet elimination on CSR / residual

Pinned by (tests/test_oracle.py): partition of unity and exact
gradients of the trilinear basis; the closed-form unit-cube stiffness
``make_KE_3d`` of feax/solvers/cmg.py:45-74 after the corner permutation;
symmetry, 6 rigid-body modes and linearity in E; central differences of the
residual (exact for linear materials); the Neo-Hookean tangent vs central
differences of P; the constant-traction closed form t*area/4 per face node;
global equilibrium sum(f_ext) = t*area.

The Gauss point order inside an element (basix's) is unpinned; it only
changes the rounding of sums over quadrature points.
"""

import numpy as np

# Reference coordinates of the 8 local nodes in feax HEX8 order: basix vertex
# v = x + 2y + 4z re-ordered by [0,1,3,2,4,5,7,6] (feax/basis.py:65-66).
XI_NODES = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                     (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.float64)

_G = (1.0 - 1.0 / np.sqrt(3.0)) / 2.0, (1.0 + 1.0 / np.sqrt(3.0)) / 2.0
# Tensor Gauss points on [0,1]^3, xi_0 slowest (basix make_quadrature).
QUAD_POINTS = np.array([(a, b, c) for a in _G for b in _G for c in _G])
QUAD_WEIGHTS = np.full(8, 1.0 / 8.0)

# basix hexahedron facets and their outward normals, facet vertices in feax
# local numbering (see oracle.structure.HEX8_FACE_INDS).
_BASIX_VERTS = np.array([(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)], float)
_BASIX_FACETS = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6), (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]
FACE_NORMALS = np.array([(0, 0, -1), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], float)
_FACE_GAUSS = np.array([(a, b) for a in _G for b in _G])
FACE_WEIGHTS = np.full((6, 4), 0.25)


def shape_vals_grads(xi):
    """N (M, 8) and dN/dxi (M, 8, 3) of the trilinear basis at points xi (M, 3)."""
    xi = np.atleast_2d(xi)
    c = XI_NODES[None, :, :]
    f = np.where(c == 1, xi[:, None, :], 1.0 - xi[:, None, :])      # (M, 8, 3)
    df = np.where(c == 1, 1.0, -1.0) * np.ones_like(f)
    N = f.prod(axis=2)
    dN = np.stack([df[..., 0] * f[..., 1] * f[..., 2],
                   f[..., 0] * df[..., 1] * f[..., 2],
                   f[..., 0] * f[..., 1] * df[..., 2]], axis=-1)
    return N, dN


SHAPE_VALS, SHAPE_GRADS_REF = shape_vals_grads(QUAD_POINTS)


def _face_quad_points():
    pts = []
    for facet in _BASIX_FACETS:
        V = _BASIX_VERTS[list(facet)]
        s, t = _FACE_GAUSS[:, 0], _FACE_GAUSS[:, 1]
        lag = np.stack([(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t], axis=1)
        pts.append(lag @ V)
    return np.stack(pts)                                             # (6, 4, 3)


FACE_QUAD_POINTS = _face_quad_points()
FACE_SHAPE_VALS, FACE_SHAPE_GRADS_REF = (
    a.reshape(6, 4, *a.shape[1:]) for a in shape_vals_grads(FACE_QUAD_POINTS.reshape(-1, 3)))


def geometry(points, cells):
    """Physical shape gradients (E, 8q, 8a, 3) and JxW (E, 8q) -- fe.py:96-129."""
    X = points[cells]                                                # (E, 8, 3)
    Jm = np.einsum("eai,qaj->eqij", X, SHAPE_GRADS_REF)              # dx_i/dxi_j
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)
    grads = np.einsum("qaj,eqji->eqai", SHAPE_GRADS_REF, inv)
    return grads, det * QUAD_WEIGHTS[None, :]


def face_geometry(points, cells, boundary_inds):
    """Face shape values (F, 4, 8) and nanson scale (F, 4) -- fe.py:175-222."""
    X = points[cells[boundary_inds[:, 0]]]                           # (F, 8, 3)
    fidx = boundary_inds[:, 1]
    g = FACE_SHAPE_GRADS_REF[fidx]                                   # (F, 4, 8, 3)
    Jm = np.einsum("fai,fqaj->fqij", X, g)
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)
    n = FACE_NORMALS[fidx]
    nanson = np.linalg.norm(np.einsum("fj,fqji->fqi", n, inv), axis=-1)
    return FACE_SHAPE_VALS[fidx], nanson * det * FACE_WEIGHTS[fidx]


# --------------------------------------------------------------------------
# Materials (stress and its tangent at quadrature points, vectorised)
# --------------------------------------------------------------------------

def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def linear_elastic_stress(ug, E, nu):
    """sigma = lambda tr(eps) I + 2 mu eps (examples/benchmark/
    bench_linear_elasticity.py:327-339); ug (..., 3, 3), E (...)."""
    lam, mu = lame(E, nu)
    eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
    tr = np.trace(eps, axis1=-2, axis2=-1)
    return lam[..., None, None] * tr[..., None, None] * np.eye(3) + 2 * mu[..., None, None] * eps


def linear_elastic_tangent(E, nu):
    """C_idjl = lam d_id d_jl + mu (d_ij d_dl + d_il d_dj), shape (..., 3,3,3,3)."""
    lam, mu = lame(E, nu)
    d = np.eye(3)
    c_lam = np.einsum("id,jl->idjl", d, d)
    c_mu = np.einsum("ij,dl->idjl", d, d) + np.einsum("il,dj->idjl", d, d)
    return lam[..., None, None, None, None] * c_lam + mu[..., None, None, None, None] * c_mu


def neo_hookean_stress(ug, mu, kappa):
    """P = mu (F - F^-T) + kappa (J-1) J F^-T (tests/test_newton_solver.py:65-89)."""
    F = np.eye(3) + ug
    J = np.linalg.det(F)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    return mu[..., None, None] * (F - FiT) + (kappa * (J - 1) * J)[..., None, None] * FiT


def neo_hookean_tangent(ug, mu, kappa):
    """dP_iJ/dF_kL of neo_hookean_stress."""
    F = np.eye(3) + ug
    J = np.linalg.det(F)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    d = np.eye(3)
    c = -mu + kappa * J * (J - 1)
    A = (mu[..., None, None, None, None] * np.einsum("ik,JL->iJkL", d, d)
         + (kappa * (2 * J - 1) * J)[..., None, None, None, None]
         * np.einsum("...iJ,...kL->...iJkL", FiT, FiT)
         - c[..., None, None, None, None] * np.einsum("...iL,...kJ->...iJkL", FiT, FiT))
    return A


def neo_hookean_energy(ug, mu, kappa):
    """psi = mu/2 (I1 - 3) - mu ln J + kappa/2 (J - 1)^2."""
    F = np.eye(3) + ug
    J = np.linalg.det(F)
    I1 = np.einsum("...ij,...ij->...", F, F)
    return 0.5 * mu * (I1 - 3) - mu * np.log(J) + 0.5 * kappa * (J - 1) ** 2


class Material:
    """Material law + where its parameter lives.

    kind: 'linear_elastic' (param = E, node or cell var), 'simp' (param = rho,
    cell var, E = Emin + rho^p (E0 - Emin)), 'neo_hookean' (param = mu, cell
    var, bulk modulus kappa).
    """

    def __init__(self, kind, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, kappa=10.0):
        self.kind, self.nu, self.E0, self.Emin, self.penal, self.kappa = kind, nu, E0, Emin, penal, kappa

    def modulus(self, p):
        if self.kind == "simp":
            return self.Emin + p ** self.penal * (self.E0 - self.Emin)
        return p

    def stress(self, ug, p):
        if self.kind == "neo_hookean":
            return neo_hookean_stress(ug, p, self.kappa)
        return linear_elastic_stress(ug, self.modulus(p), self.nu)

    def tangent(self, ug, p):
        if self.kind == "neo_hookean":
            return neo_hookean_tangent(ug, p, self.kappa)
        return linear_elastic_tangent(self.modulus(p), self.nu)


def param_at_quads(param, cells, num_nodes):
    """Interpolate a volume var to quads (E, 8q) -- assembler.py:64-103, 227-269:
    node var -> N . v_e, cell var -> broadcast."""
    param = np.asarray(param, float)
    if param.ndim == 0:
        return np.full((len(cells), 8), float(param))
    if param.shape[0] == num_nodes:
        return np.einsum("qa,ea->eq", SHAPE_VALS, param[cells])
    if param.shape[0] == len(cells):
        return np.repeat(param[:, None], 8, axis=1)
    raise ValueError(f"unsupported volume var shape {param.shape}")


# --------------------------------------------------------------------------
# Element kernels
# --------------------------------------------------------------------------

def element_residual(u_e, grads, JxW, pq, mat):
    """(E, 24) volume residual -- assembler.py:406-448."""
    ug = np.einsum("eai,eqad->eqid", u_e, grads)
    sig = mat.stress(ug, pq)
    return np.einsum("eqid,eqad,eq->eai", sig, grads, JxW).reshape(len(u_e), 24)


def element_stiffness(u_e, grads, JxW, pq, mat):
    """(E, 24, 24) element Jacobian jac[a*3+i, b*3+j] -- assembler.py:808-818."""
    n = len(u_e)
    if mat.kind == "neo_hookean":
        ug = np.einsum("eai,eqad->eqid", u_e, grads)
        C = mat.tangent(ug, pq)
        K = np.einsum("eqad,eqidjl,eqbl,eq->eaibj", grads, C, grads, JxW)
        return K.reshape(n, 24, 24)
    lam, mu = lame(mat.modulus(pq), mat.nu)
    wl, wm = lam * JxW, mu * JxW
    K = (np.einsum("eqai,eqbj,eq->eaibj", grads, grads, wl)
         + np.einsum("eqaj,eqbi,eq->eaibj", grads, grads, wm))
    gg = np.einsum("eqad,eqbd,eq->eab", grads, grads, wm)
    K += gg[:, :, None, :, None] * np.eye(3)[None, None, :, None, :]
    return K.reshape(n, 24, 24)


def face_residual(face_vals, nanson, traction):
    """(F, 24) surface residual for a dead traction t (F, 4, 3) -- +sign,
    assembler.py:494-535."""
    return np.einsum("fqi,fqa,fq->fai", traction, face_vals, nanson).reshape(len(face_vals), 24)


# --------------------------------------------------------------------------
# Global assembly
# --------------------------------------------------------------------------

class Model:
    """Everything the oracle needs for one HEX8 vec=3 problem."""

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
            self.csr = st.csr_structure(If, Jf, self.n, I, J, num_cells=self.E)
            self.nse = self.csr["nse"]
            self.vol_len = self.E * 576

    # residual R(u) = volume + surface (traction: list of (F, 4, 3) arrays)
    def residual(self, u, param, mat, tractions=()):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        vals = [element_residual(ue, self.grads, self.JxW, pq, mat).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        for (fv, ns), b, t in zip(self.faces, self.boundary_inds_list, tractions):
            vals.append(face_residual(fv, ns, t).reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return np.bincount(np.concatenate(dofs).astype(np.int64), np.concatenate(vals),
                           minlength=self.n)

    def element_matrices(self, u, param, mat, chunk=8192):
        ue = u.reshape(-1, 3)[self.cells]
        pq = param_at_quads(param, self.cells, self.N)
        out = np.empty((self.E, 24, 24))
        for s in range(0, self.E, chunk):
            e = slice(s, s + chunk)
            out[e] = element_stiffness(ue[e], self.grads[e], self.JxW[e], pq[e], mat)
        return out

    def csr_values(self, u, param, mat):
        """CSR data (nse,) summed in ascending element order (faces add
        nothing for dead loads) -- assembler.py:1253-1318, 1362-1404."""
        Ke = self.element_matrices(u, param, mat)
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        data = np.bincount(slots, Ke.reshape(-1), minlength=self.nse + 1)
        return data[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))


def apply_bc_csr(data, csr, bc_mask, bc_rows, symmetric=True):
    """Zero BC rows (and cols), set BC diagonal slots to 1 -- DCboundary.py:147-199."""
    zero = bc_mask[csr["row_of_slot"]]
    if symmetric:
        zero = zero | bc_mask[csr["indices"]]
    out = np.where(zero, 0.0, data)
    out[csr["diag_slots"][bc_rows]] = 1.0
    return out


def apply_bc_res(res, sol, bc_rows, bc_vals, scale=1.0):
    """res[bc] = sol[bc] - bc_vals*scale -- DCboundary.py:202-251."""
    out = res.copy()
    out[bc_rows] = sol[bc_rows] - bc_vals * scale
    return out


def constant_traction(boundary_inds, magnitude, direction=(0.0, 0.0, 1.0)):
    """(F, 4, 3) traction = direction * surface var (F, 4) -- the surface map
    ``lambda u, x, t: [0, 0, t]`` of tests/conftest.py:78 and the benchmark."""
    m = np.broadcast_to(np.asarray(magnitude, float), (len(boundary_inds), 4))
    return m[:, :, None] * np.asarray(direction, float)[None, None, :]
