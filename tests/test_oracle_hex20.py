"""Pins of the HEX20 oracle (oracle/hex20.py): the basis is basix's
serendipity degree-2 hexahedron (vertex values and edge means as degrees of
freedom) in feax's node order, and the element kernels on it."""

import numpy as np
import pytest

from oracle import hex8, hex20, structure as st

G3 = np.array([0.5 - np.sqrt(0.15), 0.5, 0.5 + np.sqrt(0.15)])    # 3-point Gauss on [0, 1]
W3 = np.array([5.0, 8.0, 5.0]) / 18.0


def _dofs(f):
    """Degrees of freedom of a field f(xi): values at the 8 vertices, means
    over the 12 edges (3-point Gauss, exact for polynomials of degree 5)."""
    out = [f(hex20.XI_NODES[v][None])[0] for v in range(8)]
    for a, b in hex20.EDGES:
        pa, pb = hex20.XI_NODES[a], hex20.XI_NODES[b]
        pts = pa[None] + G3[:, None] * (pb - pa)[None]
        out.append(float(W3 @ f(pts)))
    return np.array(out)


def test_dual_basis_and_partition_of_unity():
    for a in range(20):
        d = _dofs(lambda x: hex20.shape_vals_grads(x)[0][:, a])
        e = np.zeros(20)
        e[a] = 1.0
        assert np.allclose(d, e, atol=1e-14), a
    rng = np.random.default_rng(0)
    x = rng.random((50, 3))
    phi, dphi = hex20.shape_vals_grads(x)
    assert np.allclose(phi.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(dphi.sum(axis=1), 0.0, atol=1e-13)


def test_quadratics_are_reproduced():
    rng = np.random.default_rng(1)
    x = rng.random((40, 3))
    phi, dphi = hex20.shape_vals_grads(x)
    mons = [lambda p: np.ones(len(p)), lambda p: p[:, 0], lambda p: p[:, 2], lambda p: p[:, 0] * p[:, 1],
            lambda p: p[:, 1] ** 2, lambda p: p[:, 2] * p[:, 0] + 3 * p[:, 1] ** 2 - p[:, 2],
            lambda p: p[:, 0] * p[:, 1] * p[:, 2], lambda p: p[:, 0] ** 2 * p[:, 1]]
    for f in mons:
        d = _dofs(f)
        assert np.allclose(phi @ d, f(x), atol=1e-13)
    # gradient of x0^2 x1 + x2^2 from its DOFs
    f = lambda p: p[:, 0] ** 2 * p[:, 1] + p[:, 2] ** 2
    g = np.einsum("mad,a->md", dphi, _dofs(f))
    ref = np.stack([2 * x[:, 0] * x[:, 1], x[:, 0] ** 2, 2 * x[:, 2]], axis=1)
    assert np.allclose(g, ref, atol=1e-12)


def test_gradients_match_central_differences():
    rng = np.random.default_rng(2)
    x = rng.random((10, 3))
    _, d = hex20.shape_vals_grads(x)
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd = (hex20.shape_vals_grads(x + e)[0] - hex20.shape_vals_grads(x - e)[0]) / (2 * h)
        assert np.allclose(d[:, :, k], fd, atol=1e-8)


def test_box_mesh_and_face_lists():
    pts, cells = hex20.box((3, 2, 2), (1.0, 2.0, 0.5))
    nx, ny, nz = 3, 2, 2
    nodes8 = (nx + 1) * (ny + 1) * (nz + 1)
    edges = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    assert len(pts) == nodes8 + edges and cells.shape == (12, 20)
    # every cell's edge nodes sit at the midpoints of its corner pairs; det J > 0
    X = pts[cells]
    assert np.allclose(X[:, 8:], 0.5 * (X[:, hex20.EDGES[:, 0]] + X[:, hex20.EDGES[:, 1]]))
    _, JxW = hex20.geometry(pts, cells)
    assert np.all(JxW > 0) and np.isclose(JxW.sum(), 3 * 4 * 1.0)
    for f, corners in enumerate(hex20.FACE_CORNERS):
        assert np.array_equal(corners, st.HEX8_FACE_INDS[f])
        # the face's 8 nodes: the only ones whose functions do not vanish on it
        vals = np.abs(hex20.FACE_SHAPE_VALS[f]).max(axis=0)
        assert set(np.flatnonzero(vals > 1e-12)) == set(hex20.FACE_NODES[f])


def _elastic_model(dims, distort=0.0, seed=0):
    pts, cells = hex20.box(dims)
    if distort:
        rng = np.random.default_rng(seed)
        inner = np.all((pts > 1e-9) & (pts < np.array(dims) - 1e-9), axis=1)
        # move corner nodes of the interior, keep edge nodes at edge midpoints
        corner_ids = np.unique(cells[:, :8])
        dp = np.zeros_like(pts)
        sel = corner_ids[inner[corner_ids]]
        dp[sel] = rng.uniform(-distort, distort, (len(sel), 3))
        pts = pts + dp
        X = pts[cells]
        mids = 0.5 * (X[:, hex20.EDGES[:, 0]] + X[:, hex20.EDGES[:, 1]])
        pts[cells[:, 8:].reshape(-1)] = mids.reshape(-1, 3)
    return hex20.Model(pts, cells)


def test_element_stiffness_symmetry_and_rigid_modes():
    m = _elastic_model((2, 2, 1), distort=0.1, seed=3)
    mat = hex8.Material("linear_elastic", nu=0.3)
    Ke = m.element_matrices(np.zeros(m.n), np.full(m.E, 2.0), mat)
    assert np.allclose(Ke, np.swapaxes(Ke, 1, 2), atol=1e-12 * np.abs(Ke).max())
    X = m.points[m.cells]                                  # (E, 20, 3)
    for e in range(m.E):
        c = X[e]
        modes = [np.tile(np.eye(3)[k], 20) for k in range(3)]
        for i, j in ((0, 1), (1, 2), (0, 2)):
            r = np.zeros((20, 3))
            r[:, i] = -c[:, j]
            r[:, j] = c[:, i]
            modes.append(r.reshape(-1))
        for v in modes:
            assert np.abs(Ke[e] @ v).max() <= 1e-11 * np.abs(Ke[e]).max() * np.abs(v).max()


def test_linear_patch_and_homogeneous_strain_energy():
    """A linear field u = A x + c (edge DOFs = edge means = midpoint values on
    straight edges) leaves zero force at interior nodes of a distorted mesh,
    and u^T K u / 2 equals vol * eps:sigma / 2."""
    m = _elastic_model((3, 3, 3), distort=0.15, seed=4)
    mat = hex8.Material("linear_elastic", nu=0.3)
    rng = np.random.default_rng(5)
    A = 1e-3 * rng.standard_normal((3, 3))
    u = (m.points @ A.T + 1e-3 * rng.standard_normal(3)).reshape(-1)
    E = np.full(m.N, 3.0)
    r = m.residual(u, E, mat)
    inner = np.all((m.points > 1e-9) & (m.points < 3 - 1e-9), axis=1)
    assert np.abs(r.reshape(-1, 3)[inner]).max() <= 1e-13
    K = m.scipy_csr(m.csr_values(np.zeros(m.n), E, mat))
    eps = 0.5 * (A + A.T)
    lam, mu = hex8.lame(3.0, 0.3)
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    vol = m.JxW.sum()
    assert np.isclose(0.5 * u @ (K @ u), 0.5 * vol * np.sum(eps * sig), rtol=1e-12)
    assert np.isclose(vol, 27.0, rtol=1e-13)


def test_uniform_traction_consistent_loads():
    """Uniform traction t on a flat face of area A: corners get -A t / 4,
    edge nodes A t / 2 (the integrals of the edge-mean basis on the face)."""
    pts, cells = hex20.box((1, 1, 1), (2.0, 3.0, 1.0))
    faces = hex20.boundary_faces(pts, cells, lambda p: np.isclose(p[0], 2.0))
    assert len(faces) == 1
    m = hex20.Model(pts, cells, [faces])
    mat = hex8.Material("linear_elastic", nu=0.3)
    t = 0.7
    r = m.residual(np.zeros(m.n), np.full(m.E, 1.0), mat, [hex8.constant_traction(faces, t)])
    rz = r.reshape(-1, 3)[:, 2]
    A = 3.0
    fnodes = cells[0, hex20.FACE_NODES[faces[0, 1]]]
    assert np.allclose(rz[fnodes[:4]], -A * t / 4, atol=1e-14)
    assert np.allclose(rz[fnodes[4:]], A * t / 2, atol=1e-14)
    assert np.isclose(rz.sum(), A * t) and np.abs(np.delete(rz, fnodes)).max() < 1e-15


def test_neo_hookean_tangent_matches_residual_differences():
    m = _elastic_model((1, 1, 2), distort=0.0)
    mat = hex8.Material("neo_hookean", kappa=10.0)
    rng = np.random.default_rng(6)
    mu = rng.uniform(0.5, 1.5, m.E)
    u = 0.03 * rng.standard_normal(m.n)
    K = m.scipy_csr(m.csr_values(u, mu, mat)).toarray()
    v = rng.standard_normal(m.n)
    h = 1e-6
    fd = (m.residual(u + h * v, mu, mat) - m.residual(u - h * v, mu, mat)) / (2 * h)
    assert np.allclose(K @ v, fd, atol=1e-7 * np.abs(fd).max())
