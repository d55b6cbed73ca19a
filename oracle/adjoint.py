"""Oracle: reverse-mode pieces of the HEX8 residual (FP64 numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:
  feax/solvers/linear.py:426-481   f_bwd: lambda = J_bc^-T v, grads = -(dR_bc/dp)^T lambda
  feax/assembler.py:1833-1858      R_bc, whose jax.vjp feax takes
  feax/gene/responses.py:10-105    compliance C = sum_faces sum_q (-t) . u nanson
  feax/solvers/cmg.py:312-323      SIMP compliance sensitivity
                                   -p rho^(p-1) (E0 - Emin) u_e^T K0_e u_e
  feax/assembler.py:1866-1961      strain energy Pi = sum_e sum_q psi JxW

Pinned by (tests/test_oracle_adjoint.py): central differences of
lambda . R(u; p) in p (cell, node and traction variables, linear, SIMP and
Neo-Hookean laws); the compliance closed form through the element matrices
(a different route from the stress contraction used here); central
differences of the compliance of the solved state; the strain energy
gradient in u equals the Neo-Hookean residual, and its mu-partial equals
central differences in mu.
"""

import numpy as np

from .hex8 import SHAPE_VALS, element_stiffness, face_geometry, lame, param_at_quads


def _field_grads(model, v):
    """(E, 8q, 3, 3) gradients d v_i / d x_d at the Gauss points."""
    ve = v.reshape(-1, 3)[model.cells]
    return np.einsum("eai,eqad->eqid", ve, model.grads)


def _mask(lam, bc_mask):
    if bc_mask is None:
        return lam
    return np.where(bc_mask, 0.0, lam)


def quad_weights_vjp(model, u, lam, param, mat, bc_mask=None):
    """w (E, 8q) = dR_e/dp_q . lambda_e at each Gauss point (chain rule of the
    material law included, shape-function weights of node variables not)."""
    ug = _field_grads(model, u)
    lg = _field_grads(model, _mask(lam, bc_mask))
    pq = param_at_quads(param, model.cells, model.N)
    if mat.kind == "neo_hookean":
        F = np.eye(3) + ug
        FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
        dens = np.einsum("eqij,eqij->eq", F - FiT, lg)
    else:
        lam0, mu0 = lame(np.ones_like(pq), mat.nu)
        eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
        tr = np.trace(eps, axis1=-2, axis2=-1)
        sig1 = lam0[..., None, None] * tr[..., None, None] * np.eye(3) + 2 * mu0[..., None, None] * eps
        dens = np.einsum("eqij,eqij->eq", sig1, lg)
        if mat.kind == "simp":
            dens = dens * mat.penal * pq ** (mat.penal - 1) * (mat.E0 - mat.Emin)
    return dens * model.JxW


def param_vjp(model, u, lam, param, mat, bc_mask=None):
    """(dR/dp)^T lambda for a cell (E,) or node (N,) volume variable."""
    w = quad_weights_vjp(model, u, lam, param, mat, bc_mask)
    param = np.asarray(param, float)
    if param.shape[0] == model.E:
        return w.sum(axis=1)
    if param.shape[0] == model.N:
        contrib = np.einsum("qa,eq->ea", SHAPE_VALS, w)
        return np.bincount(model.cells.reshape(-1), contrib.reshape(-1), minlength=model.N)
    raise ValueError("param must be a cell or node variable")


def compliance_sens_closed_form(model, u, rho, mat):
    """-p rho_e^(p-1) (E0 - Emin) u_e^T K0_e u_e with K0_e the unit-modulus
    element stiffness (feax/solvers/cmg.py:312-323)."""
    from .hex8 import Material
    unit = Material("linear_elastic", nu=mat.nu)
    ue = u.reshape(-1, 3)[model.cells].reshape(model.E, 24)
    K0 = element_stiffness(np.zeros((model.E, 8, 3)), model.grads, model.JxW, np.ones((model.E, 8)), unit)
    quad = np.einsum("ei,eij,ej->e", ue, K0, ue)
    return -mat.penal * rho ** (mat.penal - 1) * (mat.E0 - mat.Emin) * quad


def traction_vjp(model, boundary_inds, direction, lam, bc_mask=None):
    """d(lambda . R)/d svar (F, 4) for the dead traction t = direction * svar."""
    fv, nanson = face_geometry(model.points, model.cells, boundary_inds)
    le = _mask(lam, bc_mask).reshape(-1, 3)[model.cells[boundary_inds[:, 0]]]   # (F, 8, 3)
    dl = le @ np.asarray(direction, float)                                      # (F, 8)
    return np.einsum("fqa,fa->fq", fv, dl) * nanson


def compliance(model, u, boundary_inds_list, tractions):
    """C = sum_faces sum_q (-t) . u nanson (feax/gene/responses.py:78-103)."""
    ue = u.reshape(-1, 3)
    total = 0.0
    for b, t in zip(boundary_inds_list, tractions):
        fv, nanson = face_geometry(model.points, model.cells, b)
        uf = np.einsum("fqa,fai->fqi", fv, ue[model.cells[b[:, 0]]])
        total += np.sum(-t * uf * nanson[:, :, None])
    return total


def neo_hookean_energy_total(model, u, mu, kappa):
    """(Pi, dPi/dmu_e) for a cell variable mu."""
    ug = _field_grads(model, u)
    F = np.eye(3) + ug
    J = np.linalg.det(F)
    I1 = np.einsum("eqij,eqij->eq", F, F)
    mq = param_at_quads(mu, model.cells, model.N)
    dmu = (0.5 * (I1 - 3) - np.log(J)) * model.JxW
    psi = mq * dmu + 0.5 * kappa * (J - 1) ** 2 * model.JxW
    return psi.sum(), dmu.sum(axis=1)


__all__ = ["quad_weights_vjp", "param_vjp", "compliance_sens_closed_form", "traction_vjp", "compliance",
           "neo_hookean_energy_total"]
