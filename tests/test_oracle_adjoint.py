"""Pins of the oracle's reverse-mode functions (oracle/adjoint.py).

Each check uses a route independent of the function under test: central
differences of the oracle residual / energy / solved compliance, the
element-matrix closed form of feax/solvers/cmg.py:312-323, and the
identity dPi/du = R for a hyperelastic law.
"""

import numpy as np
import pytest

from oracle import adjoint, hex8, structure


def _model(dims=(3, 2, 2), seed=1, faces_at=None):
    pts, cells = structure.structured_grid(dims)
    rng = np.random.default_rng(seed)
    pts = pts + rng.uniform(-0.12, 0.12, pts.shape)
    binds = []
    if faces_at is not None:
        L = float(dims[0])
        pts0, _ = structure.structured_grid(dims)
        binds = [structure.boundary_faces(pts0, cells, lambda p: np.isclose(p[0], L))]
    return hex8.Model(pts, cells, binds, structure=False), rng


def _fd(f, p, idx, eps):
    out = []
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        out.append((f(pp) - f(pm)) / (2 * eps))
    return np.array(out)


@pytest.mark.parametrize("kind,loc", [("linear_elastic", "cell"), ("linear_elastic", "node"),
                                      ("simp", "cell"), ("simp", "node"), ("neo_hookean", "cell")])
def test_param_vjp_matches_central_differences(kind, loc):
    m, rng = _model()
    mat = hex8.Material(kind, nu=0.3, E0=1.0, Emin=1e-3, penal=3.0, kappa=10.0)
    size = m.E if loc == "cell" else m.N
    p = rng.uniform(0.4, 1.0, size)
    scale = 0.05 if kind == "neo_hookean" else 1.0
    u = scale * rng.standard_normal(m.n)
    lam = rng.standard_normal(m.n)
    mask = np.zeros(m.n, bool)
    mask[rng.choice(m.n, 7, replace=False)] = True
    g = adjoint.param_vjp(m, u, lam, p, mat, bc_mask=mask)
    lam_m = np.where(mask, 0.0, lam)
    idx = rng.choice(size, 6, replace=False)
    fd = _fd(lambda q: lam_m @ m.residual(u, q, mat), p, idx, 1e-6)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-7, atol=1e-9 * np.abs(g).max())


def test_compliance_sensitivity_closed_form():
    """lambda = u in the stress route == -p rho^(p-1)(E0-Emin) u^T K0 u."""
    m, rng = _model((4, 3, 2), seed=3)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    rho = rng.uniform(0.3, 1.0, m.E)
    u = rng.standard_normal(m.n)
    g = -adjoint.param_vjp(m, u, u, rho, mat)
    cf = adjoint.compliance_sens_closed_form(m, u, rho, mat)
    np.testing.assert_allclose(g, cf, rtol=1e-11, atol=1e-14 * np.abs(cf).max())
    # a mistake that would pass the line above: dropping the rho^(p-1) factor
    assert not np.allclose(cf, cf / (rho ** 2), rtol=1e-3)


def test_compliance_gradient_of_solved_state():
    """dC/drho of the clamped, loaded, solved cantilever vs central differences
    of C(rho) = f . u(rho) with fresh direct solves (self-adjoint: lambda = u)."""
    import scipy.sparse.linalg as spla
    dims = (4, 2, 2)
    m, rng = _model(dims, seed=4, faces_at=True)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    rho = rng.uniform(0.3, 1.0, m.E)
    pts0, _ = structure.structured_grid(dims)
    bc_rows, _ = structure.dirichlet_rows(pts0, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    mask = np.zeros(m.n, bool)
    mask[bc_rows] = True
    t = [hex8.constant_traction(m.boundary_inds_list[0], 0.25)]
    mfull = hex8.Model(m.points, m.cells, m.boundary_inds_list)

    def solve(r):
        K = mfull.scipy_csr(hex8.apply_bc_csr(mfull.csr_values(np.zeros(m.n), r, mat), mfull.csr, mask, bc_rows))
        f = -mfull.residual(np.zeros(m.n), r, mat, t)
        f[bc_rows] = 0.0
        return spla.spsolve(K.tocsc(), f)

    def C(r):
        return adjoint.compliance(m, solve(r), m.boundary_inds_list, t)

    u = solve(rho)
    assert C(rho) > 0
    g = adjoint.compliance_sens_closed_form(m, u, rho, mat)
    idx = rng.choice(m.E, 5, replace=False)
    fd = _fd(C, rho, idx, 1e-5)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6)


def test_traction_vjp_matches_central_differences():
    m, rng = _model((3, 2, 2), seed=5, faces_at=True)
    mat = hex8.Material("linear_elastic", nu=0.3)
    b = m.boundary_inds_list[0]
    d = np.array([0.3, -0.2, 1.0])
    lam = rng.standard_normal(m.n)
    u = rng.standard_normal(m.n)
    E = np.ones(m.E)
    g = adjoint.traction_vjp(m, b, d, lam)
    s0 = rng.uniform(0.5, 1.5, (len(b), 4))
    flat = s0.reshape(-1)
    idx = np.arange(flat.size)

    def f(s):
        return lam @ m.residual(u, E, mat, [s.reshape(-1, 4)[:, :, None] * d])
    np.testing.assert_allclose(g.reshape(-1), _fd(f, flat, idx, 1e-3), rtol=1e-9, atol=1e-12)


def test_neo_hookean_energy_gradients():
    m, rng = _model((2, 2, 2), seed=6)
    mat = hex8.Material("neo_hookean", kappa=10.0)
    mu = rng.uniform(0.5, 1.5, m.E)
    u = 0.05 * rng.standard_normal(m.n)
    Pi, dmu = adjoint.neo_hookean_energy_total(m, u, mu, mat.kappa)
    # dPi/du == R(u) (psi and the stress law agree)
    R = m.residual(u, mu, mat)
    idx = rng.choice(m.n, 8, replace=False)
    fd_u = _fd(lambda v: adjoint.neo_hookean_energy_total(m, v, mu, mat.kappa)[0], u, idx, 1e-6)
    np.testing.assert_allclose(R[idx], fd_u, rtol=1e-6, atol=1e-9)
    # partial in mu
    fd_mu = _fd(lambda q: adjoint.neo_hookean_energy_total(m, u, q, mat.kappa)[0], mu, range(m.E), 1e-6)
    np.testing.assert_allclose(dmu, fd_mu, rtol=1e-7, atol=1e-12)
    # closed form at u = 0: psi = 0
    assert abs(adjoint.neo_hookean_energy_total(m, np.zeros(m.n), mu, mat.kappa)[0]) < 1e-14
