"""Oracle: a monolithic two-variable (u, T) thermo-elastic HEX8 problem (FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates feax's multi-variable assembly for Problem(mesh=[m, m], vec=[3, 1])
(feax/problem.py:155-212: per-element DOFs = variable 0's 3 c + i then
variable 1's 3N + c, offset[1] = 3N; feax/assembler.py:591-647: the weak
form's grad terms integrated against each variable's test gradients; the
residual flattened variable by variable) for the weak form
    grad_terms = [sigma, k grad T],  mass_terms = [0, 0],
    sigma = E (lam0 tr eps I + 2 mu0 eps) - E alpha (T - t_ref) / (1 - 2 nu) I,
with E a node or cell variable, constants nu, alpha, t_ref, k; dead
surface loads t e on variable 0 and a flux q on variable 1 (+ sign, feax's
surface kernels, feax/assembler.py:494-535).  The Jacobian is the element
derivative in feax's DOF order (K_Tu = 0: the heat equation does not see
u), assembled by the argsort CSR structure of oracle/structure.py and the
Dirichlet elimination of oracle/hex8.py.

Pinned by (tests/test_oracle_coupled.py): its blocks equal the HEX8
elasticity matrix (K_uu), the scalar conduction matrix (K_TT, oracle
conduction.py) and zero (K_Tu); K_uT equals central differences of the
residual in T (exact: the residual is linear); the monolithic solution
equals the staggered one (conduction solve, then the thermo-elastic solve
of oracle/thermoelastic.py with that temperature).
"""

import numpy as np

from . import hex8
from . import structure as st


class Model:
    def __init__(self, points, cells, boundary_inds_list=()):
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.E = len(self.cells)
        self.N = len(self.points)
        self.n = 4 * self.N
        self.boundary_inds_list = [np.asarray(b, np.int64).reshape(-1, 2) for b in boundary_inds_list]
        self.grads, self.JxW = hex8.geometry(self.points, self.cells)
        self.faces = [hex8.face_geometry(self.points, self.cells, b) for b in self.boundary_inds_list]
        c = self.cells
        u_dofs = (3 * c[:, :, None] + np.arange(3)[None, None, :]).reshape(self.E, 24)
        self.inds = np.concatenate([u_dofs, 3 * self.N + c], axis=1).astype(np.int64)       # (E, 32)
        I = np.repeat(self.inds[:, :, None], 32, axis=2).reshape(-1)
        J = np.repeat(self.inds[:, None, :], 32, axis=1).reshape(-1)
        for b in self.boundary_inds_list:
            fi = self.inds[b[:, 0]]
            I = np.concatenate([I, np.repeat(fi[:, :, None], 32, axis=2).reshape(-1)])
            J = np.concatenate([J, np.repeat(fi[:, None, :], 32, axis=1).reshape(-1)])
        self.csr = st.csr_structure(I, J, self.n, num_cells=self.E, ndof_vol=32)
        self.nse = self.csr["nse"]

    def _split(self, w):
        u = w[:3 * self.N].reshape(-1, 3)[self.cells]          # (E, 8, 3)
        T = w[3 * self.N:][self.cells]                          # (E, 8)
        return u, T

    def element_residual(self, w, E, law):
        nu, alpha, t_ref, k = law
        ue, Te = self._split(w)
        Eq = hex8.param_at_quads(E, self.cells, self.N)
        lam0, mu0 = hex8.lame(1.0, nu)
        ug = np.einsum("eai,eqad->eqid", ue, self.grads)
        eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
        Tq = np.einsum("qa,ea->eq", hex8.SHAPE_VALS, Te)
        tr = np.trace(eps, axis1=-2, axis2=-1)
        sig = Eq[..., None, None] * (lam0 * tr[..., None, None] * np.eye(3) + 2 * mu0 * eps)
        sig -= (Eq * alpha * (Tq - t_ref) / (1.0 - 2.0 * nu))[..., None, None] * np.eye(3)
        ru = np.einsum("eqid,eqad,eq->eai", sig, self.grads, self.JxW).reshape(self.E, 24)
        gT = np.einsum("ea,eqad->eqd", Te, self.grads)
        rT = k * np.einsum("eqd,eqad,eq->ea", gT, self.grads, self.JxW)
        return np.concatenate([ru, rT], axis=1)

    def face_residual(self, loads):
        """loads: per boundary (variable_index, direction (3,) or None, (F, 4) values)."""
        vals, dofs = [], []
        for (fv, ns), b, (var, d, s) in zip(self.faces, self.boundary_inds_list, loads):
            r = np.zeros((len(b), 32))
            if var == 0:
                r[:, :24] = np.einsum("fq,fqa,fq,i->fai", s, fv, ns, np.asarray(d, float)).reshape(len(b), 24)
            else:
                r[:, 24:] = np.einsum("fq,fqa,fq->fa", s, fv, ns)
            vals.append(r.reshape(-1))
            dofs.append(self.inds[b[:, 0]].reshape(-1))
        return vals, dofs

    def residual(self, w, E, law, loads=()):
        vals = [self.element_residual(w, E, law).reshape(-1)]
        dofs = [self.inds.reshape(-1)]
        fv, fd = self.face_residual(loads)
        return np.bincount(np.concatenate(dofs + fd), np.concatenate(vals + fv), minlength=self.n)

    def element_matrices(self, E, law):
        """(E, 32, 32) derivative of the element residual (linear in w: exact
        from the residual of unit vectors)."""
        nu, alpha, t_ref, k = law
        Eq = hex8.param_at_quads(E, self.cells, self.N)
        mat = hex8.Material("linear_elastic", nu=nu)
        Kuu = hex8.element_stiffness(np.zeros((self.E, 8, 3)), self.grads, self.JxW, Eq, mat)
        # d r_(a,i) / d T_b = -sum_q E alpha / (1 - 2 nu) N_b dN_a/dx_i JxW
        KuT = -np.einsum("eq,qb,eqai,eq->eaib", Eq * alpha / (1.0 - 2.0 * nu), hex8.SHAPE_VALS, self.grads,
                         self.JxW).reshape(self.E, 24, 8)
        KTT = k * np.einsum("eqad,eqbd,eq->eab", self.grads, self.grads, self.JxW)
        K = np.zeros((self.E, 32, 32))
        K[:, :24, :24] = Kuu
        K[:, :24, 24:] = KuT
        K[:, 24:, 24:] = KTT
        return K

    def csr_values(self, E, law):
        Ke = self.element_matrices(E, law)
        slots = self.csr["volume_slots"].reshape(-1).astype(np.int64)
        return np.bincount(slots, Ke.reshape(-1), minlength=self.nse + 1)[:self.nse]

    def scipy_csr(self, data):
        import scipy.sparse as sp
        return sp.csr_matrix((data, self.csr["indices"], self.csr["indptr"]), shape=(self.n, self.n))


def bc_rows(points, specs):
    """(rows, vals) of Dirichlet specs (location, variable_index, component,
    value) on the (u, T) layout: variable 0 rows 3 node + c, variable 1 rows
    3N + node; sorted, first occurrence kept (feax/DCboundary.py:415-494)."""
    N = len(points)
    rows, vals = [], []
    for loc, var, comp, value in specs:
        nodes = np.flatnonzero(st._location_per_node(points, loc))
        comps = ([0, 1, 2] if comp == "all" else [int(comp)]) if var == 0 else [0]
        for c in comps:
            rows.append(3 * nodes + c if var == 0 else 3 * N + nodes)
            vals.append(np.array([float(value(p)) if callable(value) else float(value) for p in points[nodes]]))
    r = np.concatenate(rows)
    v = np.concatenate(vals)
    s = np.argsort(r, kind="stable")
    r, v = r[s], v[s]
    ur, ui = np.unique(r, return_index=True)
    return ur.astype(np.int32), v[ui]
