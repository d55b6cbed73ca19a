"""Oracle: Newton with Armijo backtracking on the FP64 oracle residual.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates feax/solvers/newton.py:809-1112: step du from J_bc du = -R_bc(u)
with the BC increments lifted (x0 of feax/solvers/common.py:23-53; here a
direct solve), Armijo on 0.5 ||R_bc||^2 (c1, rho, max_backtracks), stop on
||R_bc|| <= tol or ||R_bc|| / ||R_bc(u0)|| <= rel_tol or max_iter.

Pinned by (tests/test_oracle.py, tests/test_oracle_newton.py): the
Neo-Hookean tangent equals central differences of the residual; the
converged state has R_free = 0 and u_bc = bc_vals; with a linear law one
Newton step equals the linear direct solve; the displacement-controlled
equilibrium energy satisfies dPi/dmu = the partial derivative (central
differences of re-solved states).  The iterate sequence of feax (which uses
CG for du) is not pinned.
"""

import numpy as np

from .hex8 import apply_bc_csr


def residual_bc(model, u, param, mat, bc_rows, bc_vals, tractions=()):
    r = model.residual(u, param, mat, tractions)
    r[bc_rows] = u[bc_rows] - bc_vals
    return r


def newton(model, mat, param, bc_rows, bc_vals, u0, tractions=(), tol=1e-6, rel_tol=1e-8, max_iter=100,
           c1=1e-4, rho=0.5, max_backtracks=30, linear="direct", linear_tol=1e-13):
    """Returns (u, iterations, residual norms).  linear: "direct" (scipy
    spsolve) or "pcg" (the oracle Jacobi-PCG to relative linear_tol, for an
    SPD tangent; faster at 100k+ DOF)."""
    import scipy.sparse.linalg as spla
    from .krylov import jacobi_dinv, pcg
    mask = np.zeros(model.n, bool)
    mask[bc_rows] = True
    u = np.array(u0, float)
    r = residual_bc(model, u, param, mat, bc_rows, bc_vals, tractions)
    r0 = rn = np.linalg.norm(r)
    hist = [rn]
    it = 0
    while it < max_iter and rn > tol and rn / max(r0, 1e-30) > rel_tol:
        K = model.scipy_csr(apply_bc_csr(model.csr_values(u, param, mat), model.csr, mask, bc_rows))
        if linear == "pcg":
            du, _, _ = pcg(K, -r, None, linear_tol, 0.0, 200000, jacobi_dinv(K.diagonal()))
        else:
            du = spla.spsolve(K.tocsc(), -r)
        alpha, ok = 1.0, False
        for _ in range(max_backtracks):
            trial = u + alpha * du
            tr = residual_bc(model, trial, param, mat, bc_rows, bc_vals, tractions)
            tn = np.linalg.norm(tr)
            if np.isfinite(tn) and 0.5 * (tn ** 2 - rn ** 2) <= c1 * alpha * (-rn ** 2):
                ok = True
                break
            alpha *= rho
        if not ok:
            raise RuntimeError("oracle Newton: line search failed")
        u, r, rn = trial, tr, tn
        it += 1
        hist.append(rn)
    return u, it, hist


def linearized_predictor(model, mat, param, bc_rows, bc_vals, u_prev, tractions=(), linear_tol=1e-13):
    """u_prev + d with K(u_prev) d = -R(u_prev) on the free DOFs and d = bc_vals -
    u_prev on the Dirichlet DOFs (a first-order continuation predictor; the
    Newton solve that follows is unchanged)."""
    from .krylov import jacobi_dinv, pcg
    mask = np.zeros(model.n, bool)
    mask[bc_rows] = True
    K = model.scipy_csr(model.csr_values(u_prev, param, mat)).tocsr()
    r = model.residual(u_prev, param, mat, tractions)
    d = np.zeros(model.n)
    d[bc_rows] = bc_vals - u_prev[bc_rows]
    free = np.flatnonzero(~mask)
    Kff = K[free][:, free]
    rhs = -r[free] - K[free] @ d
    df, _, _ = pcg(Kff.tocsr(), rhs, None, linear_tol, 0.0, 200000, jacobi_dinv(Kff.diagonal()))
    d[free] = df
    return u_prev + d
