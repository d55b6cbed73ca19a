"""Pins of the oracle Newton solver (oracle/newton.py)."""

import numpy as np

from oracle import adjoint, hex8, newton, structure


def _block(dims=(3, 2, 4), seed=0):
    pts, cells = structure.structured_grid(dims)
    m = hex8.Model(pts, cells)
    nz = dims[2]
    bottom, _ = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[2], 0.0), "all", 0.0)])
    top, _ = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[2], nz), 2, 0.0)])
    rows = np.concatenate([bottom, top])
    order = np.argsort(rows, kind="stable")
    rows = rows[order]
    is_top = np.concatenate([np.zeros(len(bottom), bool), np.ones(len(top), bool)])[order]
    mu = np.random.default_rng(seed).uniform(0.5, 1.5, m.E)
    return m, rows, is_top, mu


def _solve_steps(m, rows, is_top, mu, mat, steps, total=-0.1):
    nz = m.points[:, 2].max()
    u = np.zeros(m.n)
    for s in range(1, steps + 1):
        vals = np.where(is_top, total * nz * s / steps, 0.0)
        u[rows] = vals
        u, it, hist = newton.newton(m, mat, mu, rows, vals, u)
    return u, vals, it, hist


def test_linear_law_one_newton_step_is_the_linear_solve():
    import scipy.sparse.linalg as spla
    m, rows, is_top, mu = _block()
    mat = hex8.Material("linear_elastic", nu=0.3)
    E = np.full(m.E, 2.0)
    vals = np.where(is_top, -0.3, 0.0)
    u0 = np.zeros(m.n)
    u0[rows] = vals
    u, it, _ = newton.newton(m, mat, E, rows, vals, u0, rel_tol=1e-12, tol=1e-13)
    assert it <= 2
    # closed form: K_ff u_f = -K_fb u_b
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), E, mat)).tocsr()
    free = np.setdiff1d(np.arange(m.n), rows)
    ref = np.zeros(m.n)
    ref[rows] = vals
    ref[free] = spla.spsolve(K[free][:, free].tocsc(), -K[free][:, rows] @ vals)
    np.testing.assert_allclose(u, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_neo_hookean_compression_converges_to_equilibrium():
    m, rows, is_top, mu = _block()
    mat = hex8.Material("neo_hookean", kappa=10.0)
    u, vals, it, hist = _solve_steps(m, rows, is_top, mu, mat, steps=2)
    r = m.residual(u, mu, mat)
    r[rows] = 0.0
    assert np.abs(r).max() < 1e-10
    np.testing.assert_array_equal(u[rows], vals)
    # quadratic convergence in the last steps
    assert hist[-1] < 1e-8 * hist[0]


def test_energy_derivative_at_equilibrium_is_the_partial():
    """Displacement control, no external load: dPi/dmu = dPi/dmu|_u (R_free = 0,
    du_bc/dmu = 0) -- the A21 identity, checked by re-solving."""
    m, rows, is_top, mu = _block((2, 2, 3), seed=1)
    mat = hex8.Material("neo_hookean", kappa=10.0)
    u, vals, _, _ = _solve_steps(m, rows, is_top, mu, mat, steps=2)
    _, partial = adjoint.neo_hookean_energy_total(m, u, mu, 10.0)
    eps = 1e-5
    for e in (0, 5, 11):
        vals_pm = []
        for sgn in (1, -1):
            mu2 = mu.copy()
            mu2[e] += sgn * eps
            u2, _, _ = newton.newton(m, mat, mu2, rows, vals, u, rel_tol=1e-14, tol=1e-13)
            vals_pm.append(adjoint.neo_hookean_energy_total(m, u2, mu2, 10.0)[0])
        fd = (vals_pm[0] - vals_pm[1]) / (2 * eps)
        assert abs(fd - partial[e]) <= 1e-6 * abs(partial[e]) + 1e-12
