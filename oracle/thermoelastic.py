"""Oracle: thermal strain of the HEX8 linear / SIMP laws (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates the thermo-elastic tensor map of feax's
examples/basic/thermo_mech_coupling.py:95-103 and tests/test_params_bridge.py:
136-143, sigma = C(E) : (eps - alpha (T - t_ref) I), on the HEX8 kernels of
oracle/hex8.py: the temperature is a node variable interpolated at the
Gauss points, and its part of the residual is
R_th[a, i] = -sum_q E_q (3 lam0 + 2 mu0) alpha (T_q - t_ref) dN_a/dx_i JxW_q
(3 lam0 + 2 mu0 = 1 / (1 - 2 nu)).

Pinned by (tests/test_oracle_thermoelastic.py): free expansion u = alpha dT x
of a body with a uniform temperature rise is stress-free (total residual 0
at every node); the clamped block's reaction on a face is
-E alpha dT / (1 - 2 nu) * area (the pressure 3 K alpha dT).
"""

import numpy as np

from . import hex8


def thermal_residual(model, T, param, mat, alpha, t_ref=0.0):
    """(n,) thermal part of the residual for node temperatures T."""
    Tq = hex8.param_at_quads(T, model.cells, model.N)           # (E, 8) interpolated with N
    E = mat.modulus(hex8.param_at_quads(param, model.cells, model.N))
    kth = E * alpha * (Tq - t_ref) / (1.0 - 2.0 * mat.nu)
    vals = -np.einsum("eq,eqai,eq->eai", kth, model.grads, model.JxW).reshape(-1)
    return np.bincount(model.inds.reshape(-1).astype(np.int64), vals, minlength=model.n)


def residual(model, u, param, mat, T, alpha, t_ref=0.0, tractions=()):
    return model.residual(u, param, mat, tractions) + thermal_residual(model, T, param, mat, alpha, t_ref)
