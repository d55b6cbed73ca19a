"""Pins of the FP64 oracle (CPU): known answers of the reference's own tests,
closed forms and invariants that a wrong sign, index or dropped term breaks."""

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle import hex8, krylov, structure as st


def fixture_mesh():
    """The reference test fixture: 10x10x10 box, mesh_size 5 -> 2x2x2 HEX8
    (tests/conftest.py:36-41), with the traction face x = 10."""
    pts, cells = st.structured_grid((2, 2, 2), (5.0, 5.0, 5.0))
    bi = st.boundary_faces(pts, cells, lambda p: np.isclose(p[0], 10.0, 1e-5))
    return pts, cells, bi


# ------------------------------------------------------------------ structure KATs
def test_known_answers_of_reference_matrix_view_tests():
    """tests/test_matrix_view.py:107-133, 174-175, 199-200."""
    pts, cells, bi = fixture_mesh()
    I, J, _ = st.coo_ij(cells, [bi])
    assert len(I) == len(J) == 6912
    for view in ("UPPER", "LOWER"):
        If, Jf, idx = st.filter_view(I, J, view)
        assert len(If) == 3600
        assert np.all(Jf >= If) if view == "UPPER" else np.all(Jf <= If)
    cs = st.csr_structure(I, J, 81)
    assert len(cs["indptr"]) == 82          # shape (81, 81)
    assert cs["nse"] == 3087 == st.structured_nse((2, 2, 2))


@pytest.mark.parametrize("dims", [(2, 2, 2), (7, 3, 4), (50, 5, 5)])
def test_nse_and_coo_length_closed_forms(dims):
    pts, cells = st.structured_grid(dims)
    bi = st.boundary_faces(pts, cells, lambda p: np.isclose(p[0], dims[0]))
    I, J, res = st.coo_ij(cells, [bi])
    assert len(bi) == dims[1] * dims[2]
    assert len(I) == 576 * (len(cells) + len(bi))
    assert len(res) == 24 * (len(cells) + len(bi))
    cs = st.csr_structure(I, J, 3 * len(pts))
    assert cs["nse"] == st.structured_nse(dims)
    assert cs["structurally_symmetric"]
    assert np.all(cs["diag_slots"] >= 0)


def test_csr_pattern_brute_force_dense():
    """Pattern == nonzero set of a dense boolean adjacency built element by
    element, on a mesh with shuffled node numbers."""
    rng = np.random.default_rng(0)
    pts, cells = st.structured_grid((3, 2, 2))
    perm = rng.permutation(len(pts))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    cells = inv[cells]
    pts = pts[perm]
    I, J, _ = st.coo_ij(cells, [])
    n = 3 * len(pts)
    cs = st.csr_structure(I, J, n)
    dense = np.zeros((n, n), bool)
    for e in cells:
        d = (3 * e[:, None] + np.arange(3)).reshape(-1)
        dense[np.ix_(d, d)] = True
    r, c = np.nonzero(dense)
    assert np.array_equal(cs["indptr"], np.concatenate([[0], np.cumsum(dense.sum(1))]))
    assert np.array_equal(cs["indices"], c)
    # raw-entry slot map points at the entry's own (row, col)
    assert np.all(cs["row_of_slot"][cs["slot"]] == I)
    assert np.all(cs["indices"][cs["slot"]] == J)


def test_hex8_numbering_matches_spgrid_conventions():
    """feax/spgrid.py:28-41: cell (cx*ny+cy)*nz+cz, node (i*(ny+1)+j)*(nz+1)+k."""
    nx, ny, nz = 4, 3, 2
    pts, cells = st.structured_grid((nx, ny, nz), (2.0, 1.0, 0.5))
    e = (2 * ny + 1) * nz + 1                        # cell (2, 1, 1)
    base = (2 * (ny + 1) + 1) * (nz + 1) + 1         # node (2, 1, 1)
    assert cells[e, 2] == base                       # local node 2 is the (0,0,0) corner
    assert np.allclose(pts[cells[e, 6]] - pts[cells[e, 2]], [2.0, 0.0, 0.0])
    assert np.allclose(pts[cells[e, 0]] - pts[cells[e, 2]], [0.0, 1.0, 0.5])


def test_dirichlet_rows_clamped_face():
    pts, cells = st.structured_grid((4, 3, 2))
    rows, vals = st.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                                         (lambda p: np.isclose(p[1], 0.0), 2, 1.5)])
    face0 = 3 * 4 * 3                                # x = 0: 3 (ny+1)(nz+1)
    assert len(rows) == face0 + 5 * 3 - 3            # y = 0 z-comp, minus overlap with x = 0
    assert np.all(np.diff(rows) > 0)
    # duplicates keep their first occurrence (the x = 0 spec, value 0)
    n0 = np.nonzero(np.isclose(pts[:, 0], 0) & np.isclose(pts[:, 1], 0))[0]
    assert np.all(vals[np.searchsorted(rows, 3 * n0 + 2)] == 0.0)


# ------------------------------------------------------------------ basis and element
def test_basis_partition_of_unity_and_nodal_interpolation():
    N, dN = hex8.shape_vals_grads(hex8.QUAD_POINTS)
    assert np.allclose(N.sum(1), 1.0)
    assert np.allclose(dN.sum(1), 0.0)
    Nn, _ = hex8.shape_vals_grads(hex8.XI_NODES)
    assert np.allclose(Nn, np.eye(8))
    assert np.isclose(hex8.QUAD_WEIGHTS.sum(), 1.0)
    # 2-point Gauss is exact for cubics on [0,1]
    assert np.isclose((hex8.QUAD_WEIGHTS * hex8.QUAD_POINTS[:, 0] ** 3).sum(), 0.25)


def make_KE_3d(nu=0.3, E=1.0):
    """Unit-cube HEX8 stiffness of feax/solvers/cmg.py:45-74: reference cell
    [-1,1]^3, corners (dx,dy,dz) with dz fastest, Voigt B and D."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    corners = np.array([(dx, dy, dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)], float)
    s = 2 * corners - 1
    g = 1 / np.sqrt(3)
    KE = np.zeros((24, 24))
    for xi, eta, ze in [(a, b, c) for a in (-g, g) for b in (-g, g) for c in (-g, g)]:
        dN = np.stack([0.125 * s[:, 0] * (1 + s[:, 1] * eta) * (1 + s[:, 2] * ze),
                       0.125 * s[:, 1] * (1 + s[:, 0] * xi) * (1 + s[:, 2] * ze),
                       0.125 * s[:, 2] * (1 + s[:, 0] * xi) * (1 + s[:, 1] * eta)])
        Jm = dN @ corners
        dNx = np.linalg.solve(Jm, dN)
        B = np.zeros((6, 24))
        B[0, 0::3], B[1, 1::3], B[2, 2::3] = dNx
        B[3, 0::3], B[3, 1::3] = dNx[1], dNx[0]
        B[4, 1::3], B[4, 2::3] = dNx[2], dNx[1]
        B[5, 0::3], B[5, 2::3] = dNx[2], dNx[0]
        KE += B.T @ D @ B * np.linalg.det(Jm)
    return KE


def unit_cube_K(E=1.0, nu=0.3):
    p1, c1 = st.structured_grid((1, 1, 1))
    m = hex8.Model(p1, c1, structure=False)
    return m.element_matrices(np.zeros(24), np.full(8, E), hex8.Material("linear_elastic", nu=nu))[0]


def test_element_stiffness_equals_closed_form_unit_cube():
    K = unit_cube_K()
    perm = st.HEX8_OFFSETS @ np.array([4, 2, 1])     # feax local node -> cmg corner id
    P = (3 * perm[:, None] + np.arange(3)).reshape(-1)
    assert np.abs(K - make_KE_3d()[np.ix_(P, P)]).max() < 1e-15


def test_element_stiffness_invariants():
    K = unit_cube_K(E=1.0)
    assert np.allclose(K, K.T, atol=1e-15)
    w = np.linalg.eigvalsh(K)
    assert w.min() > -1e-12
    assert (np.abs(w) < 1e-10).sum() == 6            # exactly the 6 rigid-body modes
    for comp in range(3):
        u = np.zeros(24)
        u[comp::3] = 1.0
        assert np.abs(K @ u).max() < 1e-14
    assert np.allclose(unit_cube_K(E=7.5), 7.5 * K)


def test_residual_linear_in_u_and_jacobian_is_its_derivative():
    """For a linear material R(u) - R(0) = K u exactly, so K's columns are
    differences of the residual (a transposed or mis-indexed K fails)."""
    rng = np.random.default_rng(1)
    pts, cells = st.structured_grid((3, 2, 2), (1.0, 0.8, 1.3))
    pts = pts + rng.uniform(-0.1, 0.1, pts.shape)
    m = hex8.Model(pts, cells)
    mat = hex8.Material("linear_elastic", nu=0.27)
    E = rng.uniform(1.0, 2.0, m.N)
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), E, mat)).toarray()
    r0 = m.residual(np.zeros(m.n), E, mat)
    u = rng.standard_normal(m.n)
    assert np.allclose(m.residual(u, E, mat) - r0, K @ u, atol=1e-12 * np.abs(K).max())
    assert np.allclose(K, K.T, atol=1e-13 * np.abs(K).max())


def test_simp_cell_var_matches_scaled_stiffness():
    pts, cells = st.structured_grid((2, 1, 1))
    m = hex8.Model(pts, cells)
    rho = np.array([0.5, 1.0])
    Ks = m.element_matrices(np.zeros(m.n), rho, hex8.Material("simp", nu=0.3, E0=2.0, Emin=1e-9, penal=3.0))
    K1 = m.element_matrices(np.zeros(m.n), np.ones(2), hex8.Material("linear_elastic", nu=0.3))
    Eref = 1e-9 + rho ** 3 * (2.0 - 1e-9)
    assert np.allclose(Ks, Eref[:, None, None] * K1, rtol=1e-14, atol=0)


def test_neo_hookean_tangent_is_derivative_of_stress():
    rng = np.random.default_rng(2)
    ug = 0.1 * rng.standard_normal((3, 3))
    mu, kappa = np.array(1.3), 10.0
    A = hex8.neo_hookean_tangent(ug, mu, kappa)
    h = 1e-6
    for k in range(3):
        for L in range(3):
            d = np.zeros((3, 3))
            d[k, L] = h
            fd = (hex8.neo_hookean_stress(ug + d, mu, kappa) - hex8.neo_hookean_stress(ug - d, mu, kappa)) / (2 * h)
            assert np.allclose(A[:, :, k, L], fd, atol=1e-7)
    # P is the derivative of psi
    for k in range(3):
        for L in range(3):
            d = np.zeros((3, 3))
            d[k, L] = h
            fd = (hex8.neo_hookean_energy(ug + d, mu, kappa) - hex8.neo_hookean_energy(ug - d, mu, kappa)) / (2 * h)
            assert np.isclose(hex8.neo_hookean_stress(ug, mu, kappa)[k, L], fd, atol=1e-7)


def test_surface_traction_closed_form():
    """Each node of a loaded bilinear face gets t * area / 4 (+ sign); the
    residual's total is t times the loaded area (feax/assembler.py:529-533)."""
    pts, cells, bi = fixture_mesh()
    m = hex8.Model(pts, cells, [bi])
    fv, ns = m.faces[0]
    assert np.allclose(ns, 25.0 / 4)                 # face area 25, 4 Gauss points
    t = hex8.constant_traction(bi, 1e-3)
    f = m.residual(np.zeros(m.n), np.zeros(m.N), hex8.Material("linear_elastic"), [t])
    fz = f.reshape(-1, 3)[:, 2]
    assert np.isclose(fz.sum(), 1e-3 * 100.0)
    right = np.isclose(pts[:, 0], 10.0)
    corner = right & np.isclose(pts[:, 1] % 10, 0) & np.isclose(pts[:, 2] % 10, 0)
    assert np.allclose(fz[corner], 1e-3 * 25.0 / 4)  # one face
    assert np.allclose(fz[right & ~corner & ~(np.isclose(pts[:, 1], 5) & np.isclose(pts[:, 2], 5))],
                       1e-3 * 25.0 / 2)              # edge nodes: two faces
    assert np.isclose(fz[right & np.isclose(pts[:, 1], 5) & np.isclose(pts[:, 2], 5)], 1e-3 * 25.0)
    assert np.allclose(fz[~right], 0.0)


def test_geometry_of_structured_cell():
    """detJ = hx hy hz > 0 with feax's corner order; JxW = detJ / 8."""
    pts, cells = st.structured_grid((1, 1, 1), (2.0, 3.0, 0.5))
    _, JxW = hex8.geometry(pts, cells)
    assert np.allclose(JxW, 3.0 / 8)


# ------------------------------------------------------------------ Krylov
def test_pcg_matches_direct_and_stopping_rule():
    pts, cells, bi = fixture_mesh()
    m = hex8.Model(pts, cells, [bi])
    rows, vals = st.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    mat = hex8.Material("linear_elastic", nu=0.3)
    u, info = krylov.linear_solve(m, mat, np.full(m.N, 70e3), [hex8.constant_traction(bi, 1e-3)], rows, vals,
                                  tol=1e-12, atol=0.0)
    ud = spl.spsolve(info["K"].tocsc(), info["b"])
    assert np.linalg.norm(u - ud) <= 1e-10 * np.linalg.norm(ud)
    assert info["rs"] <= (1e-12) ** 2 * (info["b"] @ info["b"])
    assert np.all(u[rows] == 0.0)
    K, b = info["K"], info["b"]
    x, k, _ = krylov.pcg(K, b, np.ones(len(b)), 1e-12, 0.0, 0)
    assert k == 0 and np.all(x == 1.0)
    _, k, _ = krylov.pcg(K, b, None, 1e-30, 0.0, 5)
    assert k == 5
    # an SPD system of size n converges in <= n steps in exact arithmetic
    A = np.diag(np.arange(1.0, 6.0))
    _, k, rs = krylov.pcg(A, np.ones(5), None, 1e-14, 0.0, 50)
    assert k <= 5 and rs < 1e-26


def test_x0_lift_and_check_convergence():
    sol = np.array([1.0, 2.0, 3.0])
    x0 = krylov.x0_lift(sol, np.array([0, 2]), np.array([5.0, 1.0]))
    assert np.array_equal(x0, [4.0, 0.0, -2.0])
    A = np.eye(3)
    assert np.all(np.isnan(krylov.check_convergence(A, np.zeros(3), np.ones(3), 0.1)))
    assert np.all(krylov.check_convergence(A, np.ones(3), np.ones(3), 0.1) == 1.0)


def test_bicgstab_oracle_solves_nonsymmetric_systems():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from oracle import krylov
    rng = np.random.default_rng(4)
    n = 60
    A = sp.random(n, n, density=0.1, random_state=5) + sp.diags(4.0 + rng.random(n))
    A = A.tocsr()
    b = rng.standard_normal(n)
    x, k, rs = krylov.bicgstab(A, b, tol=1e-12, atol=0.0, maxiter=500, dinv=1.0 / A.diagonal())
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert rs <= (1e-12) ** 2 * (b @ b) and 0 < k < 500
    xs, info = spla.bicgstab(A, b, rtol=1e-12, atol=0.0, maxiter=500)
    assert info == 0 and np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xd)
    # A = I: one step is exact (s = 0 -> early exit)
    x1, k1, rs1 = krylov.bicgstab(sp.identity(n).tocsr(), b, tol=1e-12, atol=0.0)
    assert k1 == 1 and np.allclose(x1, b) and rs1 == 0.0
    # maxiter caps the count; maxiter = 0 returns x0
    _, k2, _ = krylov.bicgstab(A, b, tol=1e-30, atol=0.0, maxiter=3)
    assert k2 == 3
    x3, k3, _ = krylov.bicgstab(A, b, x0=np.ones(n), maxiter=0)
    assert k3 == 0 and np.all(x3 == 1.0)


def test_bicgstab_oracle_on_row_only_dirichlet_elimination():
    """symmetric_elimination=False keeps the BC columns: a non-symmetric
    operator whose solution equals the symmetric elimination's."""
    import scipy.sparse.linalg as spla
    from oracle import hex8, krylov, structure
    pts, cells = structure.structured_grid((4, 2, 2), (1.0, 1.0, 1.0))
    faces = structure.boundary_faces(pts, cells, lambda p: np.isclose(p[0], 4.0))
    m = hex8.Model(pts, cells, [faces])
    rows, vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                                                (lambda p: np.isclose(p[0], 4.0), 2, 0.01)])
    mat = hex8.Material("linear_elastic", nu=0.3)
    tr = [hex8.constant_traction(faces, 1e-3)]
    E = np.full(m.N, 70e3)
    u_s, info_s = krylov.linear_solve(m, mat, E, tr, rows, vals, tol=1e-12, atol=0.0)
    u0 = np.zeros(m.n)
    u0[rows] = vals
    mask = np.zeros(m.n, bool)
    mask[rows] = True
    data = hex8.apply_bc_csr(m.csr_values(u0, E, mat), m.csr, mask, rows, symmetric=False)
    K = m.scipy_csr(data)
    assert abs(K - K.T).max() > 0  # not symmetric
    b = -hex8.apply_bc_res(m.residual(u0, E, mat, tr), u0, rows, vals)
    x0 = krylov.x0_lift(u0, rows, vals)
    du, k, _ = krylov.bicgstab(K, b, x0, tol=1e-12, atol=0.0, dinv=krylov.jacobi_dinv(K.diagonal()))
    assert np.linalg.norm(u0 + du - u_s) <= 1e-8 * np.linalg.norm(u_s)
    assert np.linalg.norm(du - spla.spsolve(K.tocsc(), b)) <= 1e-8 * np.linalg.norm(du)


def _nonsym(n, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    A = (sp.random(n, n, density=0.15, random_state=seed) + sp.diags(3.0 + rng.random(n))).tocsr()
    return A, rng.standard_normal(n)


def test_gmres_cycle_is_the_minimal_residual_over_the_krylov_space():
    """One cycle of m steps returns x0 + the minimiser of ||M (b - A x)||
    over x0 + K_m(MA, M r0), computed here from an explicit basis."""
    A, b = _nonsym(40, 6)
    dinv = 1.0 / A.diagonal()
    x0 = np.random.default_rng(1).standard_normal(40)
    r0 = dinv * (b - A @ x0)
    u, rn = krylov._safe_normalize(r0)
    for m in (1, 3, 8):
        x, _, rn1, steps = krylov.gmres_cycle(lambda v: A @ v, lambda v: dinv * v, b, x0, u, rn, m)
        K = np.empty((40, m))
        v = r0.copy()
        for j in range(m):
            K[:, j] = v / np.linalg.norm(v)
            v = dinv * (A @ K[:, j])
        Q, _ = np.linalg.qr(K)
        B = dinv[:, None] * (A @ Q)
        y, *_ = np.linalg.lstsq(B, r0, rcond=None)
        xs = x0 + Q @ y
        assert steps == m
        assert abs(rn1 - np.linalg.norm(dinv * (b - A @ xs))) <= 1e-10 * rn
        assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)


def test_gmres_oracle_solves_and_stops_like_jax():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    A, b = _nonsym(60, 4)
    dinv = 1.0 / A.diagonal()
    xd = spla.spsolve(A.tocsc(), b)
    # restart >= n: one cycle is exact (restart is clipped to n)
    x, k, rn = krylov.gmres(A, b, tol=1e-12, atol=0.0, restart=100, dinv=dinv)
    assert k == 1 and np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)
    # restarted: converges, stops once ||M r|| <= ptol = ||M b|| tol
    x, k, rn = krylov.gmres(A, b, tol=1e-11, atol=0.0, restart=5, dinv=dinv)
    assert k > 1 and rn <= np.linalg.norm(dinv * b) * 1e-11
    assert abs(rn - np.linalg.norm(dinv * (b - A @ x))) <= 1e-14 * np.linalg.norm(dinv * b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    # ... and that cycle count is the first at which the bound holds
    xp, kp, rnp = krylov.gmres(A, b, tol=1e-11, atol=0.0, restart=5, dinv=dinv, maxiter=k - 1)
    assert kp == k - 1 and rnp > np.linalg.norm(dinv * b) * 1e-11
    # A = I: breakdown at the first step, x = b after one cycle
    xi, ki, rni = krylov.gmres(sp.identity(60).tocsr(), b, tol=1e-12, atol=0.0)
    assert ki == 1 and np.allclose(xi, b, rtol=1e-15, atol=1e-15) and rni <= 1e-15
    # maxiter = 0 returns x0; b = 0 with atol = 0 runs no cycle (NaN bound)
    x3, k3, _ = krylov.gmres(A, b, x0=np.ones(60), maxiter=0)
    assert k3 == 0 and np.all(x3 == 1.0)
    x4, k4, _ = krylov.gmres(A, np.zeros(60), x0=np.ones(60), tol=1e-10, atol=0.0)
    assert k4 == 0 and np.all(x4 == 1.0)
    # b = 0 with atol > 0: ptol = ||M b|| = 0, runs to the exact zero
    x5, k5, rn5 = krylov.gmres(A, np.zeros(60), x0=np.ones(60), tol=1e-10, atol=1e-10, restart=60)
    assert k5 >= 1 and np.abs(x5).max() < 1e-12


def test_gmres_oracle_on_row_only_dirichlet_elimination():
    import scipy.sparse.linalg as spla
    from oracle import hex8, structure
    pts, cells = structure.structured_grid((4, 2, 2), (1.0, 1.0, 1.0))
    faces = structure.boundary_faces(pts, cells, lambda p: np.isclose(p[0], 4.0))
    m = hex8.Model(pts, cells, [faces])
    rows, vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    mat = hex8.Material("linear_elastic", nu=0.3)
    E = np.full(m.N, 70e3)
    u0 = np.zeros(m.n)
    mask = np.zeros(m.n, bool)
    mask[rows] = True
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(u0, E, mat), m.csr, mask, rows, symmetric=False))
    b = -hex8.apply_bc_res(m.residual(u0, E, mat, [hex8.constant_traction(faces, 1e-3)]), u0, rows, vals)
    du, k, _ = krylov.gmres(K, b, np.zeros(m.n), tol=1e-12, atol=0.0, dinv=krylov.jacobi_dinv(K.diagonal()))
    xd = spla.spsolve(K.tocsc(), b)
    assert np.linalg.norm(du - xd) <= 1e-9 * np.linalg.norm(xd) and k > 1
