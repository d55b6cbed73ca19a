"""Pins of the multigrid oracle (oracle/gmg.py) by the mathematics that fixes
each piece; CPU only."""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gmg, hex8, structure


def _field(x):
    """A trilinear vector field (every term of Q1 on the box)."""
    return np.stack([1 + 2 * x[:, 0] - x[:, 1] + 0.5 * x[:, 0] * x[:, 1] * x[:, 2],
                     x[:, 2] - x[:, 0] * x[:, 2] + 0.25 * x[:, 1],
                     3 + x[:, 1] * x[:, 2] - x[:, 0] * x[:, 1]], axis=1)


@pytest.mark.parametrize("dims", [(4, 2, 6), (4, 3, 2), (2, 2, 2)])
def test_prolongation_reproduces_trilinear_fields(dims):
    c = gmg.coarse_dims(dims, 1)
    h = [a // b for a, b in zip(dims, c)]
    pf, _ = structure.structured_grid(dims)
    pc, _ = structure.structured_grid(c, h)
    P = gmg.prolongation(dims, c)
    assert np.abs(P @ _field(pc).reshape(-1) - _field(pf).reshape(-1)).max() < 1e-13
    # injection picks the fine node with the same coordinates
    inj = gmg.node_injection(dims, c)
    assert np.array_equal(pf[inj], pc)


@pytest.mark.parametrize("dims", [(5, 3, 3), (7, 1, 3)])
def test_ceil_halving_prolongation_reproduces_trilinear_fields(dims):
    """Odd counts are ceil-halved: the coarse grid (spacing 2) reaches one
    fine cell past the fine one and P still reproduces trilinear fields."""
    c = gmg.coarse_dims(dims, 1)
    assert all(b == (a + 1) // 2 or b == a for a, b in zip(dims, c)) and c != dims
    h = [1 if a == b else 2 for a, b in zip(dims, c)]
    pf, _ = structure.structured_grid(dims)
    pc, _ = structure.structured_grid(c, h)
    P = gmg.prolongation(dims, c)
    assert np.abs(P @ _field(pc).reshape(-1) - _field(pf).reshape(-1)).max() < 1e-12
    # children inside the fine grid only; the mean of a constant is the constant
    ch = gmg.cell_children(dims, c)
    assert np.array_equal(np.sort(ch[ch >= 0]), np.arange(int(np.prod(dims))))
    assert np.allclose(gmg.coarsen_cell_var(np.full(int(np.prod(dims)), 3.5), dims, c), 3.5)
    inj = gmg.node_injection(dims, c)
    assert inj.max() < len(pf)


@pytest.mark.parametrize("dims", [(4, 2, 2), (4, 2, 3), (2, 4, 4)])
def test_galerkin_product_equals_rediscretised_operator(dims):
    """Nested Q1 spaces with exact 2x2x2 Gauss on both grids: P^T K_h P = K_H
    for a uniform material (fixes weights, numbering and coarse assembly)."""
    c = gmg.coarse_dims(dims, 1)
    h = [a // b for a, b in zip(dims, c)]
    mat = hex8.Material("linear_elastic", nu=0.3)
    pts, cells = structure.structured_grid(dims)
    m = hex8.Model(pts, cells)
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), np.full(m.E, 7.0), mat))
    pts2, cells2 = structure.structured_grid(c, h)
    m2 = hex8.Model(pts2, cells2)
    K2 = m2.scipy_csr(m2.csr_values(np.zeros(m2.n), np.full(m2.E, 7.0), mat)).toarray()
    P = gmg.prolongation(dims, c)
    G = (P.T @ K @ P).toarray()
    assert np.abs(G - K2).max() <= 1e-13 * np.abs(K2).max()


def test_galerkin_coarsest_equals_rediscretised_for_uniform_material():
    """Uniform material, even grid: the Galerkin coarsest operator (with the
    masks and identity rows) equals the rediscretised BC-eliminated one."""
    dims = (4, 2, 2)
    pts, cells = structure.structured_grid(dims)
    mask = np.zeros(3 * len(pts), bool)
    mask[(3 * np.nonzero(pts[:, 0] == 0.0)[0][:, None] + np.arange(3)).reshape(-1)] = True
    mat = hex8.Material("linear_elastic", nu=0.3)
    levels = gmg.hierarchy(dims, np.full(len(cells), 5.0), mat, mask, nlev=2, min_cells=1)
    G = gmg.galerkin_dense(levels[0], levels[1]["mask"])
    A1 = levels[1]["A"].toarray()
    assert np.abs(G - A1).max() <= 1e-12 * np.abs(A1).max()
    assert np.allclose(levels[1]["dense_inv"], np.linalg.inv(A1), rtol=1e-9, atol=1e-12 * np.abs(np.linalg.inv(A1)).max())


def test_cell_children_and_coarsened_average():
    dims, c = (4, 2, 6), (2, 1, 3)
    ch = gmg.cell_children(dims, c)
    assert ch.shape == (6, 8) and np.array_equal(np.sort(ch.reshape(-1)), np.arange(48))
    # children lie inside their parent: centroids within the coarse cell
    pf, cf = structure.structured_grid(dims)
    pc, cc = structure.structured_grid(c, (2, 2, 2))
    cent_f = pf[cf].mean(axis=1)
    cent_c = pc[cc].mean(axis=1)
    assert np.all(np.abs(cent_f[ch] - cent_c[:, None, :]).max(axis=(1, 2)) <= 0.5 + 1e-12)


@pytest.mark.parametrize("degree", [1, 2, 3, 5])
def test_chebyshev_error_polynomial(degree):
    lam = np.linspace(0.02, 2.0, 41)
    A = np.diag(lam)
    lmax, ratio = 2.0, 10.0
    xstar = np.linspace(-1.0, 2.0, 41)
    b = lam * xstar
    x0 = np.cos(np.arange(41.0))
    hi, lo = 1.1 * lmax, lmax / ratio
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    Tk = np.zeros(degree + 1)
    Tk[-1] = 1.0
    e = np.polynomial.chebyshev.chebval((theta - lam) / delta, Tk) / np.polynomial.chebyshev.chebval(theta / delta, Tk)
    x = gmg.chebyshev(A, np.ones(41), b, None, lmax, ratio, degree, True)
    assert np.allclose(xstar - x, e * xstar, atol=1e-12)
    x = gmg.chebyshev(A, np.ones(41), b, x0, lmax, ratio, degree, False)
    assert np.allclose(xstar - x, e * (xstar - x0), atol=1e-12)


def _cantilever(dims, kind="simp", seed=0):
    rng = np.random.default_rng(seed)
    pts, cells = structure.structured_grid(dims)
    mask = np.zeros(3 * len(pts), bool)
    mask[(3 * np.nonzero(pts[:, 0] == 0.0)[0][:, None] + np.arange(3)).reshape(-1)] = True
    if kind == "simp":
        mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)
        param = rng.uniform(0.2, 1.0, len(cells))
    else:
        mat = hex8.Material("linear_elastic", nu=0.3)
        param = rng.uniform(50.0, 90.0, len(pts))
    return pts, cells, mask, mat, param


def test_vcycle_is_symmetric_positive_definite():
    dims = (4, 4, 2)
    pts, cells, mask, mat, param = _cantilever(dims)
    levels = gmg.hierarchy(dims, param, mat, mask, min_cells=1)
    assert len(levels) == 3
    n = levels[0]["A"].shape[0]
    V = np.stack([gmg.vcycle(levels, e, 3, 40, 15.0) for e in np.eye(n)], axis=1)
    assert np.abs(V - V.T).max() <= 1e-12 * np.abs(V).max()
    free = ~mask
    ev = np.linalg.eigvalsh(0.5 * (V + V.T)[np.ix_(free, free)])
    assert ev.min() > 0


def test_single_level_dense_vcycle_is_the_inverse():
    dims = (2, 2, 2)
    pts, cells, mask, mat, param = _cantilever(dims, "linear")
    levels = gmg.hierarchy(dims, param, mat, mask, nlev=1)
    A = levels[0]["A"]
    b = np.where(mask, 0.0, np.sin(np.arange(A.shape[0]) * 1.0))
    x, k = gmg.gmg_pcg(A, b, None, lambda r: gmg.vcycle(levels, r, 3, 40, 15.0), tol=1e-12)
    assert k == 1
    assert np.allclose(A @ x, b, atol=1e-10 * np.abs(b).max())


def test_gmg_pcg_with_ceil_halved_levels_matches_direct_solve():
    dims = (10, 6, 6)
    pts, cells, mask, mat, param = _cantilever(dims, "simp", seed=8)
    levels = gmg.hierarchy(dims, param, mat, mask, min_cells=1)
    assert [l["dims"] for l in levels][:3] == [(10, 6, 6), (5, 3, 3), (3, 2, 2)]
    A = levels[0]["A"]
    b = np.where(mask, 0.0, np.random.default_rng(2).standard_normal(A.shape[0]))
    x, k = gmg.gmg_pcg(A, b, None, lambda r: gmg.vcycle(levels, r, 2, 40, 8.0), tol=1e-10)
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd) and k < 40


@pytest.mark.parametrize("kind", ["simp", "linear"])
def test_gmg_pcg_matches_direct_solve(kind):
    dims = (8, 4, 4)
    pts, cells, mask, mat, param = _cantilever(dims, kind, seed=3)
    levels = gmg.hierarchy(dims, param, mat, mask, min_cells=1)
    assert [l["dims"] for l in levels] == [(8, 4, 4), (4, 2, 2), (2, 1, 1), (1, 1, 1)]
    A = levels[0]["A"]
    rng = np.random.default_rng(1)
    b = np.where(mask, 0.0, rng.standard_normal(A.shape[0]))
    x, k = gmg.gmg_pcg(A, b, None, lambda r: gmg.vcycle(levels, r, 3, 40, 15.0), tol=1e-10)
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    assert k < 40


def test_lanczos_bounds_lie_inside_and_converge_to_the_spectrum_ends():
    dims = (4, 2, 2)
    pts, cells, mask, mat, param = _cantilever(dims)
    levels = gmg.hierarchy(dims, param, mat, mask, nlev=1, dense=False)
    A, dinv = levels[0]["A"], levels[0]["dinv"]
    s = np.sqrt(dinv)
    ev = np.linalg.eigvalsh((A.toarray() * s[:, None]) * s[None, :])
    for steps in (3, 10, 30):
        lo, hi = gmg.lanczos_bounds(A, dinv, steps)
        assert ev.min() * (1 - 1e-10) <= lo <= hi <= ev.max() * (1 + 1e-10)
    lo, hi = gmg.lanczos_bounds(A, dinv, 10)
    assert hi >= 0.97 * ev.max()
    lo, hi = gmg.lanczos_bounds(A, dinv, 60)
    assert abs(hi - ev.max()) <= 1e-8 * ev.max() and abs(lo - ev.min()) <= 1e-6 * ev.max()


def test_lanczos_bounds_survive_exact_convergence():
    """A tiny operator converges before the requested steps (breakdown)."""
    A = np.diag([1.0, 2.0, 5.0])
    lo, hi = gmg.lanczos_bounds(A, np.ones(3), 10)
    assert abs(lo - 1.0) < 1e-10 and abs(hi - 5.0) < 1e-10
