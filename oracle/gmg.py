"""Oracle: geometric multigrid on a structured HEX8 box (FP64 numpy/scipy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The preconditioner of feax_amd's GeometricMultigrid, the analogue of feax's
NarrowBandCMG (feax/solvers/cmg.py:95-300): rediscretised coarse operators
(cmg.py:375-386, 426-431 average the SIMP density of the child cells),
trilinear transfers (cmg.py:187-216: weights 1 or 1/2, 1/2 per direction
for an even or odd fine index), V-cycle preconditioned CG (cmg.py:255-300).  The
smoother here is a Chebyshev polynomial in D^-1 A (cmg.py uses coloured
block Gauss-Seidel) and the coarsest level a dense inverse (cmg.py: cuDSS).
A preconditioner changes the iterates, not the solution: the product's
GMG-PCG solution is compared with the direct solve.

Pinned by (tests/test_oracle_gmg.py):
  prolongation  reproduces every trilinear field on the coarse grid exactly
                (in particular linear ones); with no constraints
                P^T K_h P = K_2h for a uniform material (nested Q1 spaces,
                2x2x2 Gauss exact on both grids), which fixes the weights,
                the node numbering and the coarse rediscretisation together
  chebyshev     on a diagonal operator the error polynomial equals
                T_k((theta - lam) / delta) / T_k(theta / delta) (numpy's
                Chebyshev series)
  vcycle        symmetric positive definite as an explicit matrix; one
                level with the exact inverse is A^-1
  galerkin_dense  equals the rediscretised coarse operator for a uniform
                material on even grids (the Galerkin identity above)
  gmg_pcg       solution equals scipy's direct solve
  lanczos_bounds  inside the spectrum of D^-1 A and converging to its ends
                (numpy eigh of the symmetrically scaled operator)
"""

import numpy as np
import scipy.sparse as sp


def coarse_dims(dims, min_cells=2):
    """Halve each direction whose cell count is even and stays >= min_cells;
    when none can be, ceil-halve the odd ones (cmg.py:82-87, 178-182: the last
    coarse cell reaches one fine cell past the grid); None when nothing can
    be coarsened."""
    out = tuple(d // 2 if (d % 2 == 0 and d // 2 >= min_cells) else d for d in dims)
    if out != tuple(dims):
        return out
    out = tuple((d + 1) // 2 if (d >= 3 and (d + 1) // 2 >= min_cells) else d for d in dims)
    return None if out == tuple(dims) else out


def prolongation_1d(nf, nc):
    """(nf + 1) x (nc + 1) node interpolation along one direction: halved
    (nf = 2 nc) -> weights 1 (even fine index) or 1/2, 1/2 (odd); kept -> I."""
    if nf == nc:
        return sp.identity(nf + 1, format="csr")
    assert nc == (nf + 1) // 2
    rows, cols, vals = [], [], []
    for i in range(nf + 1):
        if i % 2 == 0:
            rows.append(i), cols.append(i // 2), vals.append(1.0)
        else:
            rows += [i, i]
            cols += [i // 2, i // 2 + 1]
            vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf + 1, nc + 1))


def prolongation(dims_f, dims_c, free_f=None, free_c=None):
    """DOF prolongation (3 N_f x 3 N_c); node id (i (ny+1) + j)(nz+1) + k
    (feax/spgrid.py:28-41), DOF 3 node + c.  free_*: bool masks; constrained
    rows / columns are zeroed."""
    Px, Py, Pz = (prolongation_1d(f, c) for f, c in zip(dims_f, dims_c))
    P = sp.kron(sp.kron(sp.kron(Px, Py), Pz), sp.identity(3)).tocsr()
    if free_f is not None:
        P = sp.diags(free_f.astype(float)) @ P
    if free_c is not None:
        P = P @ sp.diags(free_c.astype(float))
    return P.tocsr()


def _factors(dims_f, dims_c):
    return tuple(1 if f == c else 2 for f, c in zip(dims_f, dims_c))


def node_injection(dims_f, dims_c):
    """Fine node id of every coarse node (same coordinates; a coarse node
    past the end of an odd fine direction takes the last fine node)."""
    fx, fy, fz = _factors(dims_f, dims_c)
    nyf, nzf = dims_f[1] + 1, dims_f[2] + 1
    i, j, k = np.meshgrid(np.arange(dims_c[0] + 1), np.arange(dims_c[1] + 1), np.arange(dims_c[2] + 1),
                          indexing="ij")
    I, J, K = np.minimum(fx * i, dims_f[0]), np.minimum(fy * j, dims_f[1]), np.minimum(fz * k, dims_f[2])
    return ((I * nyf + J) * nzf + K).reshape(-1)


def cell_children(dims_f, dims_c):
    """(E_c, k) fine cell ids inside every coarse cell (k = product of the
    per-direction factors), -1 where the coarse cell reaches past the fine
    grid."""
    fx, fy, fz = _factors(dims_f, dims_c)
    nx, ny, nz = dims_f
    ci, cj, ck = np.meshgrid(np.arange(dims_c[0]), np.arange(dims_c[1]), np.arange(dims_c[2]), indexing="ij")
    out = []
    for a in range(fx):
        for b in range(fy):
            for c in range(fz):
                I, J, K = fx * ci + a, fy * cj + b, fz * ck + c
                ids = (I * ny + J) * nz + K
                out.append(np.where((I < nx) & (J < ny) & (K < nz), ids, -1))
    return np.stack([o.reshape(-1) for o in out], axis=1)


def coarsen_cell_var(p, dims_f, dims_c):
    """Mean of a cell field over the child cells inside the fine grid."""
    ch = cell_children(dims_f, dims_c)
    ok = ch >= 0
    return np.where(ok, p[np.maximum(ch, 0)], 0.0).sum(axis=1) / ok.sum(axis=1)


def chebyshev(A, dinv, b, x, lmax, ratio, degree, zero_init, lmin=0.0):
    """Chebyshev iteration (Saad, Iterative Methods, Alg. 12.1) for A x = b
    with preconditioner D^-1 on [lmax / ratio, 1.1 lmax] ([lmin, 1.1 lmax]
    when lmin > 0)."""
    hi, lo = 1.1 * lmax, (lmin if lmin > 0 else lmax / ratio)
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    x = np.zeros_like(b) if zero_init else x.copy()
    r = b - A @ x
    d = dinv * r / theta
    x = x + d
    rho_prev = 1.0 / sigma
    for _ in range(1, degree):
        r = b - A @ x
        rho = 1.0 / (2.0 * sigma - rho_prev)
        d = rho * rho_prev * d + (2.0 * rho / delta) * dinv * r
        x = x + d
        rho_prev = rho
    return x


def vcycle(levels, b, smooth_degree, coarse_degree, ratio, l=0):
    """z = V(b): levels[l] = dict(A, dinv, lmax, P (to level l+1, masked),
    free, dense_inv (coarsest, optional))."""
    L = levels[l]
    if l == len(levels) - 1:
        if L.get("dense_inv") is not None:
            return (L["dense_inv"] @ b) * L["free"]
        return chebyshev(L["A"], L["dinv"], b, None, L["lmax"], ratio, coarse_degree, True, L["lmin"])
    x = chebyshev(L["A"], L["dinv"], b, None, L["lmax"], ratio, smooth_degree, True)
    r = b - L["A"] @ x
    xc = vcycle(levels, L["P"].T @ r, smooth_degree, coarse_degree, ratio, l + 1)
    x = x + L["P"] @ xc
    return chebyshev(L["A"], L["dinv"], b, x, L["lmax"], ratio, smooth_degree, False)


def lanczos_bounds(A, dinv, steps):
    """(lmin, lmax) of the Lanczos tridiagonal of D^-1 A from `steps`
    Jacobi-PCG steps on A x = v, v_i = 1 + (i mod 7) / 4 (Golub and Van Loan,
    Matrix Computations, 10.2): T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1},
    T_{j,j+1} = sqrt(beta_j)/alpha_j; steps after a breakdown are dropped."""
    n = A.shape[0]
    r = 1.0 + 0.25 * (np.arange(n) % 7)
    p = dinv * r
    gamma = r @ (dinv * r)
    alphas, betas = [], []
    for _ in range(steps):
        q = A @ p
        alpha = gamma / (p @ q)
        r = r - alpha * q
        rz = r @ (dinv * r)
        beta = rz / gamma
        gamma = rz
        alphas.append(alpha)
        betas.append(beta)
        p = dinv * r + beta * p
    m = 0
    while m < steps and np.isfinite(alphas[m]) and alphas[m] > 0 and np.isfinite(betas[m]) and betas[m] > 0:
        m += 1
    if m < steps and np.isfinite(alphas[m]) and alphas[m] > 0:
        m += 1
    a, b = np.asarray(alphas[:m]), np.asarray(betas[:m])
    T = np.diag(1.0 / a)
    T[1:, 1:] += np.diag(b[:m - 1] / a[:m - 1])
    off = np.sqrt(b[:m - 1]) / a[:m - 1]
    T += np.diag(off, 1) + np.diag(off, -1)
    ev = np.linalg.eigvalsh(T)
    return ev.min(), ev.max()


def gmg_pcg(A, b, x0, M, tol=1e-10, atol=0.0, maxiter=1000):
    """CG with preconditioner M (a function), jax.scipy.sparse.linalg.cg
    iteration and stopping rule ||r||^2 <= max(tol^2 ||b||^2, atol^2)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A @ x
    z = M(r)
    p = z.copy()
    gamma = r @ z
    atol2 = max(tol * tol * (b @ b), atol * atol)
    k = 0
    while r @ r > atol2 and k < maxiter:
        q = A @ p
        alpha = gamma / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        z = M(r)
        g_new = r @ z
        p = z + (g_new / gamma) * p
        gamma = g_new
        k += 1
    return x, k


def galerkin_dense(fine_level, coarse_mask):
    """P'^T A P' (dense) of a level's operator with its masked transfer to
    the next grid; constrained coarse DOFs get identity rows and columns."""
    P = fine_level["P"]
    G = (P.T @ fine_level["A"] @ P).toarray()
    fixed = np.asarray(coarse_mask, bool)
    G[fixed, :] = 0.0
    G[:, fixed] = 0.0
    G[fixed, fixed] = 1.0
    return G


def hierarchy(dims, param, mat, bc_mask, nlev=None, min_cells=2, lmax_iters=10, dense=True, shift=1e-12,
              spacing=(1.0, 1.0, 1.0), coarse_dofs=1):
    """Level list for vcycle(): rediscretised BC-eliminated operators on the
    grids of coarse_dims until a level has <= coarse_dofs DOFs (node var:
    value at the same node; cell var: mean of the child cells inside the
    fine grid), constrained DOFs injected from bc_mask, dinv =
    1 / (diag + shift), lmax (and the coarsest lmin) by lanczos_bounds,
    masked transfers, and with ``dense`` the inverse of the Galerkin
    product P'^T A P' of the level above as the coarsest solve."""
    from . import hex8, structure
    levels = []
    d = tuple(dims)
    h = np.asarray(spacing, float)
    p = np.asarray(param, float)
    mask = np.asarray(bc_mask, bool)
    while True:
        pts, cells = structure.structured_grid(d, h)
        model = hex8.Model(pts, cells)
        data = model.csr_values(np.zeros(model.n), p, mat)
        rows = np.nonzero(mask)[0]
        A = model.scipy_csr(hex8.apply_bc_csr(data, model.csr, mask, rows, True))
        lev = dict(A=A, dims=d, free=~mask, dinv=1.0 / (A.diagonal() + shift), param=p, mask=mask)
        levels.append(lev)
        c = coarse_dims(d, min_cells)
        if c is None or (nlev is not None and len(levels) == nlev) or model.n <= coarse_dofs:
            break
        inj = node_injection(d, c)
        if p.shape[0] == model.N:
            pc = p[inj]
        else:
            pc = coarsen_cell_var(p, d, c)
        maskc = mask.reshape(-1, 3)[inj].reshape(-1)
        lev["P"] = prolongation(d, c, ~mask, ~maskc)
        h = h * np.asarray(_factors(d, c), float)
        d, p, mask = c, pc, maskc
    for i, lev in enumerate(levels):
        last = i == len(levels) - 1
        if last and dense:
            lev["dense_inv"] = np.linalg.inv(galerkin_dense(levels[i - 1], lev["mask"]) if i > 0
                                             else lev["A"].toarray())
            lev["lmax"] = lev["lmin"] = 0.0
        else:
            lo, hi = lanczos_bounds(lev["A"], lev["dinv"], lmax_iters)
            lev["lmax"], lev["lmin"] = hi, (lo if last else 0.0)
    return levels
