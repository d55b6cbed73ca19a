"""Pins of the scalar HEX8 conduction oracle (oracle/conduction.py) by what
the physics fixes, and the scalar host structure (native vec = 1 pattern)
against it.  CPU."""

import numpy as np

from oracle import conduction
from oracle import structure as st


def _distorted(dims=(4, 3, 2), h=(1.0, 0.5, 2.0), seed=0):
    pts, cells = st.structured_grid(dims, h)
    L = np.array(dims) * np.array(h)
    inner = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    rng = np.random.default_rng(seed)
    pts = pts.copy()
    pts[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 3)) * np.array(h)
    return pts, cells, L


def test_element_matrices_have_the_constants_as_kernel():
    pts, cells, _ = _distorted()
    m = conduction.Model(pts, cells)
    K = m.element_matrices(np.linspace(0.5, 2.0, m.N))
    assert np.abs(K - K.transpose(0, 2, 1)).max() <= 1e-14 * np.abs(K).max()
    assert np.abs(K.sum(axis=2)).max() <= 1e-13 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K[0])
    assert (np.abs(ev) < 1e-12 * ev.max()).sum() == 1 and ev.min() > -1e-12 * ev.max()


def test_linear_field_patch_test_and_energy():
    """T = g.x on a distorted mesh: interior residual 0, face reactions
    k g.n area, and 1/2 T^T K T = 1/2 V k |g|^2."""
    pts, cells, L = _distorted()
    m = conduction.Model(pts, cells)
    k = 3.0
    g = np.array([0.7, -0.2, 0.4])
    T = pts @ g
    r = m.residual(T, np.full(m.E, k))
    interior = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    assert np.abs(r[interior]).max() <= 1e-13 * np.abs(r).max()
    on_x = np.isclose(pts[:, 0], L[0])
    assert abs(r[on_x].sum() - k * g[0] * L[1] * L[2]) <= 1e-12 * k * abs(g[0]) * L[1] * L[2]
    K = m.scipy_csr(m.csr_values(np.full(m.E, k)))
    W = 0.5 * np.prod(L) * k * g @ g
    assert abs(0.5 * T @ (K @ T) - W) <= 1e-12 * W


def test_prescribed_flux_total_and_bar_profile():
    """A flux q on x = L puts q * area into the residual; a bar held at T0
    and T1 on its end faces has the exact linear profile (nodal exactness
    of the linear solution, any k > 0 constant over the bar)."""
    import scipy.sparse.linalg as spla
    pts, cells, L = _distorted()
    b = st.boundary_faces(pts, cells, lambda p: np.isclose(p[0], L[0]))
    m = conduction.Model(pts, cells, [b])
    q = 0.6
    r = m.residual(np.zeros(m.n), np.ones(m.E), fluxes=[q])
    assert abs(r.sum() - q * L[1] * L[2]) <= 1e-13 * q * L[1] * L[2]
    # bar: T = 0 on x = 0, T = 100 on x = L
    K = m.scipy_csr(m.csr_values(np.full(m.E, 2.5)))
    fixed = np.isclose(pts[:, 0], 0.0) | np.isclose(pts[:, 0], L[0])
    Tb = np.where(np.isclose(pts[:, 0], L[0]), 100.0, 0.0)
    free = ~fixed
    Kff = K[free][:, free]
    rhs = -(K[free][:, fixed] @ Tb[fixed])
    T = Tb.copy()
    T[free] = spla.spsolve(Kff.tocsc(), rhs)
    assert np.abs(T - 100.0 * pts[:, 0] / L[0]).max() <= 1e-10 * 100.0


def test_simp_conductivity_law():
    rho = np.array([0.0, 0.5, 1.0])
    k = conduction.conductivity("simp", rho, k0=2.0, kmin=0.01, penal=3.0)
    assert np.allclose(k, [0.01, 0.01 + 0.125 * 1.99, 2.0])


def test_native_scalar_structure_matches_oracle():
    import feax_amd as fe
    pts, cells, L = _distorted()
    rng = np.random.default_rng(5)
    perm = rng.permutation(len(pts))
    pts, cells = pts[perm], np.argsort(perm)[cells][rng.permutation(len(cells))]

    class Thermal(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicConduction()

        def get_surface_maps(self):
            return [fe.materials.SurfaceFlux()]

    loc = lambda p: np.isclose(p[0], L[0])
    p = Thermal(fe.Mesh(pts, cells), vec=1, dim=3, location_fns=[loc])
    b = st.boundary_faces(pts, cells, loc)
    assert np.array_equal(p.boundary_inds_list[0], b)
    m = conduction.Model(pts, cells, [b])
    assert p.kernel_tag == "hex8s" and p.num_total_dofs_all_vars == len(pts) and p.csr_nse == m.nse
    assert np.array_equal(p.csr_indptr, m.csr["indptr"]) and np.array_equal(p.csr_indices, m.csr["indices"])
    assert np.array_equal(p.csr_volume_slots, m.csr["volume_slots"])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 5.0)]).create_bc(p)
    assert np.array_equal(bc.bc_rows, np.flatnonzero(np.isclose(pts[:, 0], 0.0))) and np.all(bc.bc_vals == 5.0)
    T = np.arange(len(pts), dtype=float)
    assert np.array_equal(fe.TracedParams.node_var_from_solution(p, T), T)
    tp = fe.TracedParams(volume_vars=(np.ones(len(cells)),)).replace_volume_var(0, np.zeros(len(cells)))
    assert np.all(tp.volume_vars[0] == 0.0)


def test_plane_conduction_patch_test_and_native_structure():
    """QUAD4 conduction: linear field on a distorted mesh (interior residual
    0, edge reaction k g.n length, energy), and the native quad4s pattern."""
    import feax_amd as fe
    from oracle import quad4
    pts, cells = quad4.rectangle(4, 3, 2.0, 1.5)
    inner = (pts[:, 0] > 1e-9) & (pts[:, 0] < 2.0 - 1e-9) & (pts[:, 1] > 1e-9) & (pts[:, 1] < 1.5 - 1e-9)
    pts = pts.copy()
    pts[inner] += np.random.default_rng(2).uniform(-0.1, 0.1, (inner.sum(), 2))
    m = conduction.QuadModel(pts, cells)
    k, g = 2.0, np.array([0.6, -0.3])
    T = pts @ g
    r = m.residual(T, np.full(m.E, k))
    assert np.abs(r[inner]).max() <= 1e-13 * np.abs(r).max()
    assert abs(r[np.isclose(pts[:, 0], 2.0)].sum() - k * g[0] * 1.5) <= 1e-12
    K = m.scipy_csr(m.csr_values(np.full(m.E, k)))
    assert abs(0.5 * T @ (K @ T) - 0.5 * 3.0 * k * g @ g) <= 1e-12
    p = fe.Problem(fe.Mesh(pts, cells, ele_type="QUAD4"), vec=1, dim=2, ele_type="QUAD4")
    assert p.kernel_tag == "quad4s" and p.csr_nse == m.nse
    assert np.array_equal(p.csr_indptr, m.csr["indptr"]) and np.array_equal(p.csr_indices, m.csr["indices"])
    assert np.array_equal(p.csr_volume_slots, m.csr["volume_slots"])


def test_tet_conduction_patch_test():
    from oracle import tet4
    pts, cells = tet4.box_tets((3, 2, 2), (1.0, 0.5, 2.0))
    m = conduction.TetModel(pts, cells)
    k, g = 1.5, np.array([0.2, -0.7, 0.4])
    T = pts @ g
    r = m.residual(T, np.full(m.E, k))
    L = np.array([3.0, 1.0, 4.0])
    interior = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    assert np.abs(r[interior]).max() <= 1e-13 * np.abs(r).max()
    K = m.scipy_csr(m.csr_values(np.full(m.E, k)))
    assert abs(0.5 * T @ (K @ T) - 0.5 * 12.0 * k * g @ g) <= 1e-12 * 12.0
