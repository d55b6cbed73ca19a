"""Pins of the QUAD4 oracle (oracle/quad4.py) by what the mechanics fixes, and
the host structure (mesh numbering, native CSR pattern, boundary edges)
against it.  CPU."""

import numpy as np

from oracle import hex8, quad4

NU = 0.3


def _distorted(nx=4, ny=3, lx=2.0, ly=1.5, seed=0):
    """A rectangle with its interior nodes moved (elements stay convex)."""
    pts, cells = quad4.rectangle(nx, ny, lx, ly)
    rng = np.random.default_rng(seed)
    inner = (pts[:, 0] > 1e-9) & (pts[:, 0] < lx - 1e-9) & (pts[:, 1] > 1e-9) & (pts[:, 1] < ly - 1e-9)
    h = min(lx / nx, ly / ny)
    pts = pts.copy()
    pts[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) * h
    return pts, cells


def test_rectangle_is_a_positive_conforming_mesh():
    pts, cells = _distorted()
    _, JxW = quad4.geometry(pts, cells)
    assert np.all(JxW > 0) and abs(JxW.sum() - 2.0 * 1.5) < 1e-13
    edges = np.sort(cells[:, quad4.QUAD4_FACE_INDS].reshape(-1, 2), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert set(counts.tolist()) <= {1, 2} and (counts == 1).sum() == 2 * (4 + 3)


def test_element_matrices_are_symmetric_with_rigid_body_kernel():
    pts, cells = _distorted()
    m = quad4.Model(pts, cells)
    for mat, p in ((hex8.Material("linear_elastic", nu=NU), np.linspace(1.0, 3.0, m.N)),
                   (hex8.Material("simp", nu=NU, E0=1.0, Emin=1e-3, penal=3.0), np.linspace(0.2, 1.0, m.E))):
        K = m.element_matrices(np.zeros(m.n), p, mat)
        assert np.abs(K - K.transpose(0, 2, 1)).max() <= 1e-14 * np.abs(K).max()
        X = pts[cells]
        modes = [np.broadcast_to([1.0, 0.0], X.shape), np.broadcast_to([0.0, 1.0], X.shape),
                 np.stack([-X[..., 1], X[..., 0]], axis=2)]
        for r in modes:
            assert np.abs(np.einsum("eij,ej->ei", K, r.reshape(len(cells), 8))).max() <= 1e-13 * np.abs(K).max()
        ev = np.linalg.eigvalsh(K[0])
        assert (np.abs(ev) < 1e-12 * ev.max()).sum() == 3


def test_constant_strain_patch_test_on_a_distorted_mesh():
    """u = A x: interior residual 0, edge reactions sigma n length, and
    1/2 u^T K u = 1/2 area sigma : eps (plane strain)."""
    lx, ly = 2.0, 1.5
    pts, cells = _distorted(lx=lx, ly=ly)
    m = quad4.Model(pts, cells)
    mat = hex8.Material("linear_elastic", nu=NU)
    E = 7.0
    A = np.array([[0.01, 0.004], [-0.003, -0.02]])
    u = (pts @ A.T).reshape(-1)
    r = m.residual(u, np.full(m.N, E), mat).reshape(-1, 2)
    eps = 0.5 * (A + A.T)
    lam, mu = hex8.lame(E, NU)
    sig = lam * np.trace(eps) * np.eye(2) + 2 * mu * eps
    interior = (pts[:, 0] > 1e-9) & (pts[:, 0] < lx - 1e-9) & (pts[:, 1] > 1e-9) & (pts[:, 1] < ly - 1e-9)
    assert np.abs(r[interior]).max() <= 1e-13 * np.abs(r).max()
    on_x = np.isclose(pts[:, 0], lx)
    assert np.allclose(r[on_x].sum(axis=0), sig[:, 0] * ly, rtol=1e-12, atol=1e-15)
    on_y0 = np.isclose(pts[:, 1], 0.0) & ~np.isclose(pts[:, 0], 0.0) & ~np.isclose(pts[:, 0], lx)
    assert on_y0.sum() == 3
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), np.full(m.N, E), mat))
    W = 0.5 * lx * ly * np.sum(sig * eps)
    assert abs(0.5 * u @ (K @ u) - W) <= 1e-12 * W


def test_dead_traction_splits_each_edge_length_between_its_vertices():
    lx, ly = 2.0, 1.5
    pts, cells = _distorted(lx=lx, ly=ly)
    for loc, length in ((lambda p: np.isclose(p[0], lx), ly), (lambda p: np.isclose(p[1], 0.0), lx)):
        b = quad4.boundary_faces(pts, cells, loc)
        m = quad4.Model(pts, cells, [b])
        t = 0.7
        r = m.residual(np.zeros(m.n), np.ones(m.N), hex8.Material("linear_elastic", nu=NU),
                       [quad4.constant_traction(b, t)]).reshape(-1, 2)
        assert abs(r[:, 1].sum() - t * length) <= 1e-13 * t * length and np.abs(r[:, 0]).max() == 0.0
        fv, ns = m.faces[0]
        X = pts[cells[b[:, 0]][np.arange(len(b))[:, None], quad4.QUAD4_FACE_INDS[b[:, 1]]]]
        le = np.linalg.norm(X[:, 1] - X[:, 0], axis=1)
        assert np.allclose(ns.sum(axis=1), le, rtol=1e-14)
        assert np.allclose(np.einsum("fqa,fq->fa", fv, ns).sum(axis=1), le, rtol=1e-14)


def test_neo_hookean_tangent_is_the_derivative_of_the_residual():
    pts, cells = _distorted(3, 2)
    m = quad4.Model(pts, cells)
    mat = hex8.Material("neo_hookean", kappa=10.0)
    rng = np.random.default_rng(3)
    mu = rng.uniform(0.5, 1.5, m.E)
    u = 0.05 * rng.standard_normal(m.n)
    K = m.scipy_csr(m.csr_values(u, mu, mat)).toarray()
    eps = 1e-6
    for j in rng.choice(m.n, 6, replace=False):
        du = np.zeros(m.n)
        du[j] = eps
        fd = (m.residual(u + du, mu, mat) - m.residual(u - du, mu, mat)) / (2 * eps)
        assert np.abs(fd - K[:, j]).max() <= 1e-7 * np.abs(K).max()


def test_neo_hookean_tangent_at_identity_is_linear_elasticity():
    """At F = I the Neo-Hookean tangent is the isotropic one with
    lambda = kappa and the same mu."""
    pts, cells = _distorted(3, 2)
    m = quad4.Model(pts, cells)
    mu, kappa = 0.8, 10.0
    Knh = m.element_matrices(np.zeros(m.n), np.full(m.E, mu), hex8.Material("neo_hookean", kappa=kappa))
    # isotropic law with lambda = kappa, mu: E, nu from (lambda, mu)
    nu = kappa / (2 * (kappa + mu))
    E = mu * 2 * (1 + nu)
    Kle = m.element_matrices(np.zeros(m.n), np.full(m.E, E), hex8.Material("linear_elastic", nu=nu))
    assert np.abs(Knh - Kle).max() <= 1e-12 * np.abs(Knh).max()


def test_native_rectangle_mesh_and_structure_match_oracle():
    """feax_amd's rectangle_mesh numbering, and the native pattern
    (fx_quad4_pattern), boundary edges and face incidence of a QUAD4
    Problem on a shuffled distorted mesh against the oracle."""
    import feax_amd as fe
    mesh = fe.mesh.rectangle_mesh(4, 3, domain_x=2.0, domain_y=1.5)
    pts0, cells0 = quad4.rectangle(4, 3, 2.0, 1.5)
    assert np.allclose(mesh.points, pts0) and np.array_equal(mesh.cells, cells0)
    pts, cells = _distorted(lx=2.0, ly=1.5)
    rng = np.random.default_rng(5)
    perm = rng.permutation(len(pts))
    pts, cells = pts[perm], np.argsort(perm)[cells][rng.permutation(len(cells))]

    class LE(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=NU)

    loc = lambda p: np.isclose(p[0], 2.0)
    p = LE(fe.Mesh(pts, cells, ele_type="QUAD4"), vec=2, dim=2, ele_type="QUAD4", location_fns=[loc])
    b = quad4.boundary_faces(pts, cells, loc)
    assert len(b) == 3 and np.array_equal(p.boundary_inds_list[0], b)
    m = quad4.Model(pts, cells, [b])
    assert p.csr_nse == m.nse and p.num_total_dofs_all_vars == 2 * len(pts)
    assert np.array_equal(p.csr_indptr, m.csr["indptr"]) and np.array_equal(p.csr_indices, m.csr["indices"])
    assert np.array_equal(p.csr_volume_slots, m.csr["volume_slots"])
    assert p.fes[0].num_quads == 4 and p.fes[0].num_face_quads == 2
    assert fe.TracedParams.create_uniform_surface_var(p, 0.5).shape == (len(b), 2)
