"""Pins of the TET4 oracle (oracle/tet4.py) by what the mechanics fixes, and
the host structure (native CSR pattern, boundary faces) against it.  CPU."""

import numpy as np
import pytest

from oracle import hex8, tet4

NU = 0.3


def _mesh(dims=(3, 2, 2), h=(1.0, 0.5, 2.0)):
    return tet4.box_tets(dims, h)


def test_box_tets_is_a_conforming_positive_mesh():
    pts, cells = _mesh()
    X = pts[cells]
    vol = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)) / 6
    assert np.all(vol > 0) and abs(vol.sum() - 3 * 1.0 * 2 * 0.5 * 2 * 2.0) < 1e-12
    faces = np.sort(cells[:, tet4.TET4_FACE_INDS].reshape(-1, 3), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert set(counts.tolist()) <= {1, 2}
    # boundary facets tile the box surface: 2 triangles per boundary quad
    nx, ny, nz = 3, 2, 2
    assert (counts == 1).sum() == 4 * (nx * ny + ny * nz + nx * nz)


def test_element_matrices_are_symmetric_with_rigid_body_kernel():
    pts, cells = _mesh()
    m = tet4.Model(pts, cells)
    mat = hex8.Material("linear_elastic", nu=NU)
    K = m.element_matrices(np.zeros(m.n), np.linspace(1.0, 3.0, m.E), mat)
    assert np.abs(K - K.transpose(0, 2, 1)).max() <= 1e-14 * np.abs(K).max()
    X = pts[cells]                                           # (E, 4, 3)
    modes = []
    for d in range(3):
        t = np.zeros_like(X)
        t[..., d] = 1.0
        modes.append(t)
    for w in np.eye(3):
        modes.append(np.cross(w, X))
    for r in modes:
        assert np.abs(np.einsum("eij,ej->ei", K, r.reshape(len(cells), 12))).max() <= 1e-13 * np.abs(K).max()
    # and no other zero modes: rank 6 deficiency exactly
    ev = np.linalg.eigvalsh(K[0])
    assert (np.abs(ev) < 1e-12 * ev.max()).sum() == 6


def test_constant_strain_patch_test_and_energy():
    """u = A x: interior residual 0, face reactions sigma n area, and
    1/2 u^T K u = 1/2 V sigma : eps (V the box volume)."""
    dims, h = (3, 2, 2), (1.0, 0.5, 2.0)
    pts, cells = _mesh(dims, h)
    m = tet4.Model(pts, cells)
    mat = hex8.Material("linear_elastic", nu=NU)
    E = 7.0
    A = np.array([[0.01, 0.002, -0.003], [0.004, -0.02, 0.001], [0.0, 0.003, 0.015]])
    u = (pts @ A.T).reshape(-1)
    r = m.residual(u, np.full(m.N, E), mat).reshape(-1, 3)
    L = np.array(dims) * np.array(h)
    eps = 0.5 * (A + A.T)
    lam, mu = hex8.lame(E, NU)
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    interior = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    assert np.abs(r[interior]).max() <= 1e-14 * np.abs(r).max()
    on_x = np.isclose(pts[:, 0], L[0])
    area_x = L[1] * L[2]
    assert np.allclose(r[on_x].sum(axis=0), sig[:, 0] * area_x, rtol=1e-12, atol=1e-15)
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), np.full(m.N, E), mat))
    V = np.prod(L)
    assert abs(0.5 * u @ (K @ u) - 0.5 * V * np.sum(sig * eps)) <= 1e-12 * 0.5 * V * np.sum(sig * eps)


def test_dead_traction_puts_area_over_three_on_each_face_vertex():
    dims, h = (2, 3, 2), (1.0, 0.5, 2.0)
    pts, cells = _mesh(dims, h)
    L = np.array(dims) * np.array(h)
    for axis, lf_loc in ((0, lambda p: np.isclose(p[0], L[0])), (2, lambda p: np.isclose(p[2], 0.0))):
        b = tet4.boundary_faces(pts, cells, lf_loc)
        m = tet4.Model(pts, cells, [b])
        t = 0.7
        r = m.residual(np.zeros(m.n), np.ones(m.N), hex8.Material("linear_elastic", nu=NU),
                       [tet4.constant_traction(b, t)]).reshape(-1, 3)
        area = np.prod(np.delete(L, axis))
        assert abs(r[:, 2].sum() - t * area) <= 1e-13 * t * area
        # per face: its three vertices share t * area_f equally
        X = pts[cells[b[:, 0]][np.arange(len(b))[:, None], tet4.TET4_FACE_INDS[b[:, 1]]]]
        af = 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)
        fv, ns = m.faces[0]
        assert np.allclose(ns[:, 0], af, rtol=1e-14)
        assert np.allclose(fv.sum(axis=2), 1.0) and np.allclose(fv.max(axis=2), 1 / 3)


def test_neo_hookean_tangent_is_the_derivative_of_the_residual():
    pts, cells = tet4.box_tets((2, 1, 1))
    m = tet4.Model(pts, cells)
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


def test_native_tet4_structure_matches_oracle():
    """The native pattern (fx_tet4_pattern), boundary faces and face
    incidence of a TET4 Problem against the oracle's argsort structure."""
    import feax_amd as fe
    pts, cells = tet4.box_tets((3, 2, 2), (1.0, 0.5, 2.0))
    rng = np.random.default_rng(5)
    perm = rng.permutation(len(pts))
    inv = np.argsort(perm)
    pts, cells = pts[perm], inv[cells][rng.permutation(len(cells))]

    class LE(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=NU)

    loc = lambda p: np.isclose(p[0], 3.0)
    p = LE(fe.Mesh(pts, cells, ele_type="TET4"), vec=3, dim=3, ele_type="TET4", location_fns=[loc])
    b = tet4.boundary_faces(pts, cells, loc)
    assert np.array_equal(p.boundary_inds_list[0], b)
    m = tet4.Model(pts, cells, [b])
    assert p.csr_nse == m.nse
    assert np.array_equal(p.csr_indptr, m.csr["indptr"]) and np.array_equal(p.csr_indices, m.csr["indices"])
    assert np.array_equal(p.csr_volume_slots, m.csr["volume_slots"])
    assert p.fes[0].num_quads == 1 and p.fes[0].num_face_quads == 1
    assert fe.TracedParams.create_uniform_surface_var(p, 0.5).shape == (len(b), 1)
