"""Oracle: Jacobi-preconditioned CG and the linear solve wrapper (FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:
  jax.scipy.sparse.linalg.cg (third-party; called at feax/solvers/common.py:232).
    Not vendored in the reference and JAX is absent here, so the published
    open-source algorithm is restated: r0 = b - A x0, z0 = M r0, p0 = z0,
    gamma0 = r0.z0; loop while ||r||^2 > max(tol^2 ||b||^2, atol^2) and
    k < maxiter: q = A p, alpha = gamma / (p.q), x += alpha p, r -= alpha q,
    z = M r, gamma' = r.z, beta = gamma'/gamma, p = z + beta p.  The
    recursively updated r is the one tested; no error on maxiter.
  jax.scipy.sparse.linalg.bicgstab (third-party; called at
    feax/solvers/common.py:241-252), restated from the published algorithm:
    r = b - A x0, rhat = p = q = r, rho = alpha = omega = 1; loop while
    ||r||^2 > max(tol^2 ||b||^2, atol^2), k < maxiter, no breakdown:
    rho' = rhat.r, beta = rho'/rho alpha/omega, p = r + beta (p - omega q),
    phat = M p, q = A phat, alpha = rho'/rhat.q, s = r - alpha q; if
    s.s < atol2: x += alpha phat, r = s; else shat = M s, t = A shat,
    omega = t.s/t.t, x += alpha phat + omega shat, r = s - omega t;
    breakdown when alpha = 0 or omega = 0 (after a full step).
  jax.scipy.sparse.linalg.gmres (third-party; called at
    feax/solvers/common.py:254-265 with restart 20 unless options.restart),
    restated from the published algorithm with its default
    solve_method='batched': left preconditioning, maxiter counts restart
    cycles, stop on ||M r|| <= ||M b|| min(1, max(tol ||b||, atol)/||b||);
    each cycle runs up to `restart` Arnoldi steps (one classical
    Gram-Schmidt pass, breakdown below eps ||M A v||) and solves the
    projected least squares problem through its normal equations.
  feax/solvers/common.py:86-94   M(x) = x / (diag + jacobi_shift)
  feax/solvers/common.py:23-53   x0 = lift of BC increments
  feax/solvers/common.py:376-395 check_convergence: NaN unless
                                 ||Ax - b|| / (||b|| + 1e-12) < threshold
  feax/solvers/linear.py:387-417 u = u0 + CG(J_bc, -R_bc(u0), x0)

Pinned by: a scipy.sparse.linalg.spsolve direct solve of the same system
(tests/test_oracle.py), the stopping-rule invariants (returned
residual below the threshold, iteration cap honoured, maxiter=0 returns x0),
and exact convergence in <= n steps on small SPD systems; bicgstab by the
direct solve of non-symmetric systems (the row-only Dirichlet elimination
and a random non-symmetric matrix), by agreeing with scipy's own
bicgstab's solution, and by one exact step on A = I; gmres by the
direct solve, by the exact solution after one cycle when restart >= n, by
the minimal-residual property of a cycle (its residual equals the least
squares optimum over x0 + K_m(MA, M r0) computed from an explicit Krylov
basis), by the Arnoldi relation M A V_m = V_{m+1} H^T and by a breakdown
on A = I.  The iterate trajectories of JAX's own cg, bicgstab and gmres
are parity unpinned (need an import).
"""

import numpy as np


def pcg(A, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, dinv=None):
    """Return (x, iters, ||r||^2) following jax.scipy.sparse.linalg.cg."""
    matvec = A if callable(A) else (lambda v: A @ v)
    M = (lambda v: v) if dinv is None else (lambda v: dinv * v)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    atol2 = max(tol * tol * float(b @ b), atol * atol)
    r = b - matvec(x)
    z = M(r)
    p = z.copy()
    gamma = float(r @ z)
    k = 0
    rs = float(r @ r)
    while rs > atol2 and k < maxiter:
        q = matvec(p)
        alpha = gamma / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        z = M(r)
        gamma_new = float(r @ z)
        beta = gamma_new / gamma
        p = z + beta * p
        gamma = gamma_new
        rs = float(r @ r)
        k += 1
    return x, k, rs


def bicgstab(A, b, x0=None, tol=1e-10, atol=1e-10, maxiter=10000, dinv=None):
    """Return (x, iters, ||r||^2) following jax.scipy.sparse.linalg.bicgstab
    (iters = steps taken; a breakdown ends the loop)."""
    matvec = A if callable(A) else (lambda v: A @ v)
    M = (lambda v: v) if dinv is None else (lambda v: dinv * v)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    atol2 = max(tol * tol * float(b @ b), atol * atol)
    r = b - matvec(x)
    rhat, p, q = r.copy(), r.copy(), r.copy()
    rho = alpha = omega = 1.0
    k = 0
    rs = float(r @ r)
    broke = False
    while rs > atol2 and k < maxiter and not broke:
        rho_new = float(rhat @ r)
        beta = rho_new / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = M(p)
        q = matvec(phat)
        alpha = rho_new / float(rhat @ q)
        s = r - alpha * q
        if float(s @ s) < atol2:
            x = x + alpha * phat
            r = s
            broke = alpha == 0.0
        else:
            shat = M(s)
            t = matvec(shat)
            omega = float(t @ s) / float(t @ t)
            x = x + alpha * phat + omega * shat
            r = s - omega * t
            broke = alpha == 0.0 or omega == 0.0
        rho = rho_new
        rs = float(r @ r)
        k += 1
    return x, k, rs


def _safe_normalize(v, thresh=None):
    """(v / ||v||, ||v||), or (0, 0) when ||v|| <= thresh (default: machine
    epsilon, an absolute threshold) -- jax's GMRES helper."""
    nrm = float(np.sqrt(v @ v))
    if thresh is None:
        thresh = np.finfo(np.float64).eps
    if nrm > thresh:
        return v / nrm, nrm
    return np.zeros_like(v), 0.0


def gmres_cycle(matvec, M, b, x, unit_residual, residual_norm, restart):
    """One restart cycle of jax.scipy.sparse.linalg.gmres, solve_method=
    'batched': Arnoldi on M A from the unit preconditioned residual (one
    classical Gram-Schmidt pass; a step whose orthogonalised norm is at most
    eps * ||M A v|| is a breakdown and ends the cycle), then the least
    squares problem min ||beta e0 - H^T y|| solved through its normal
    equations (H H^T) y = H beta by Cholesky, H = eye(restart, restart+1)
    with row j overwritten by step j; x += V[:, :restart] y; the new
    preconditioned residual M (b - A x) is normalised.
    Returns (x, unit_residual, residual_norm, steps)."""
    n = b.size
    eps = np.finfo(np.float64).eps
    V = np.zeros((n, restart + 1))
    V[:, 0] = unit_residual
    H = np.eye(restart, restart + 1)
    j = 0
    breakdown = False
    while j < restart and not breakdown:
        w = M(matvec(V[:, j]))
        _, w0 = _safe_normalize(w)
        h = V.T @ w
        w = w - V @ h
        unit_w, w1 = _safe_normalize(w, eps * w0)
        V[:, j + 1] = unit_w
        h[j + 1] = w1
        H[j, :] = h
        breakdown = w1 == 0.0
        j += 1
    beta = np.zeros(restart + 1)
    beta[0] = residual_norm
    a = H.T
    L = np.linalg.cholesky(a.T @ a)
    y = np.linalg.solve(L.T, np.linalg.solve(L, a.T @ beta))
    x = x + V[:, :-1] @ y
    unit, nrm = _safe_normalize(M(b - matvec(x)))
    return x, unit, nrm, j


def gmres(A, b, x0=None, tol=1e-10, atol=1e-10, restart=20, maxiter=10000, dinv=None):
    """Return (x, cycles, ||M (b - A x)||) following jax.scipy.sparse.linalg.
    gmres (solve_method='batched', its default; called at
    feax/solvers/common.py:254-265 with restart = options.restart or 20):
    left-preconditioned, restart = min(restart, n), maxiter counts restart
    cycles, and the loop runs while cycles < maxiter and the preconditioned
    residual norm exceeds ptol = ||M b|| * min(1, max(tol ||b||, atol) / ||b||)
    (NaN, so no cycle, when b = 0 and atol = 0)."""
    matvec = A if callable(A) else (lambda v: A @ v)
    M = (lambda v: v) if dinv is None else (lambda v: dinv * v)
    n = b.size
    restart = min(restart, n)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    bn = float(np.sqrt(b @ b))
    atol_eff = max(tol * bn, atol)
    Mb = M(b)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.float64(atol_eff) / np.float64(bn)
    ptol = float(np.sqrt(Mb @ Mb)) * (np.nan if np.isnan(ratio) else min(1.0, float(ratio)))
    unit, rn = _safe_normalize(M(b - matvec(x)))
    k = 0
    while k < maxiter and rn > ptol:
        x, unit, rn, _ = gmres_cycle(matvec, M, b, x, unit, rn, restart)
        k += 1
    return x, k, rn


def jacobi_dinv(diag, shift=1e-12):
    """1 / (diag + shift) -- feax/solvers/common.py:86-94."""
    return 1.0 / (diag + shift)


def x0_lift(sol, bc_rows, bc_vals):
    """x0[bc] = bc_vals - sol[bc], 0 elsewhere -- feax/solvers/common.py:23-53."""
    x0 = np.zeros_like(sol)
    x0[bc_rows] = bc_vals - sol[bc_rows]
    return x0


def check_convergence(A, x, b, threshold=0.1):
    """NaN-fill unless the relative residual < threshold -- common.py:376-395."""
    matvec = A if callable(A) else (lambda v: A @ v)
    rel = np.linalg.norm(matvec(x) - b) / (np.linalg.norm(b) + 1e-12)
    return x if rel < threshold else np.full_like(x, np.nan)


def linear_solve(model, mat, param, tractions, bc_rows, bc_vals, u0=None,
                 tol=1e-10, atol=1e-10, maxiter=10000, jacobi=True, shift=1e-12,
                 symmetric=True):
    """One linear solve u = u0 + CG(J_bc, -R_bc(u0)) on the assembled CSR.

    Returns (u, info) with info = dict(iters, rs, K (scipy, BC-applied), b, x0).
    feax/solvers/linear.py:387-417 with the CSR operator of
    feax/assembler.py:1695-1713 (identical to the matrix-free one,
    assembler.py:1753-1754).
    """
    from . import hex8
    n = model.n
    bc_mask = np.zeros(n, bool)
    bc_mask[bc_rows] = True
    if u0 is None:
        u0 = np.zeros(n)
        u0[bc_rows] = bc_vals
    data = model.csr_values(u0, param, mat)
    data_bc = hex8.apply_bc_csr(data, model.csr, bc_mask, bc_rows, symmetric)
    K = model.scipy_csr(data_bc)
    res = model.residual(u0, param, mat, tractions)
    res_bc = hex8.apply_bc_res(res, u0, bc_rows, bc_vals)
    b = -res_bc
    x0 = x0_lift(u0, bc_rows, bc_vals)
    dinv = jacobi_dinv(K.diagonal(), shift) if jacobi else None
    du, iters, rs = pcg(K, b, x0, tol, atol, maxiter, dinv)
    return u0 + du, dict(iters=iters, rs=rs, K=K, b=b, x0=x0, data=data, data_bc=data_bc)
