"""Host-side structure of the product (native fx_hex8_pattern & friends, the
Problem / DirichletBC / TracedParams mirrors) against the oracle.  These
call host-only routines of libfeaxhip.so; no GPU is touched."""

import ctypes

import numpy as np
import pytest

from oracle import structure as st


def LE():
    import feax_amd as fe

    class LinearElasticity(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]
    return LinearElasticity


def shuffled_mesh(dims, seed):
    rng = np.random.default_rng(seed)
    pts, cells = st.structured_grid(dims, (1.0, 1.3, 0.7))
    perm = rng.permutation(len(pts))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    return pts[perm], inv[cells][rng.permutation(len(cells))]


@pytest.mark.parametrize("case", ["fixture", "cantilever50x5x5", "odd", "shuffled"])
def test_native_pattern_matches_oracle_argsort(case):
    import feax_amd as fe
    if case == "fixture":
        pts, cells = st.structured_grid((2, 2, 2), (5.0, 5.0, 5.0))
        L = 10.0
    elif case == "cantilever50x5x5":
        pts, cells = st.structured_grid((50, 5, 5), (2.0, 2.0, 2.0))
        L = 100.0
    elif case == "odd":
        pts, cells = st.structured_grid((7, 3, 4))
        L = 7.0
    else:
        pts, cells = shuffled_mesh((4, 3, 3), 11)
        L = 4.0
    loc = lambda p: np.isclose(p[0], L)
    p = LE()(fe.Mesh(pts, cells), vec=3, dim=3, location_fns=[loc])
    bi = st.boundary_faces(pts, cells, loc)
    assert np.array_equal(bi, p.boundary_inds_list[0])
    I, J, _ = st.coo_ij(cells, [bi])
    cs = st.csr_structure(I, J, 3 * len(pts), num_cells=len(cells))
    assert p.csr_nse == cs["nse"]
    assert np.array_equal(p.csr_indptr, cs["indptr"])
    assert np.array_equal(p.csr_indices, cs["indices"])
    assert np.array_equal(p.csr_diag_slots, cs["diag_slots"])
    assert np.array_equal(p.csr_row_of_slot, cs["row_of_slot"])
    # elem_rank reproduces the reference's per-element slot map csr_volume_slots
    rank = p.elem_rank.reshape(-1, 8, 8).astype(np.int64)
    rows = 3 * p.cells.astype(np.int64)[:, :, None] + np.arange(3)
    slot = (p.csr_indptr[rows][:, :, :, None, None] + 3 * rank[:, :, None, :, None]
            + np.arange(3)[None, None, None, None, :])
    assert np.array_equal(slot.reshape(len(cells), 576), cs["volume_slots"])
    # node -> element incidence is complete and ascending
    for nd in range(0, len(pts), max(1, len(pts) // 50)):
        lst = p.node_inc[p.node_inc_ptr[nd]:p.node_inc_ptr[nd + 1]]
        e, a = lst >> 3, lst & 7
        assert np.all(np.diff(e) > 0)
        assert np.all(p.cells[e, a] == nd)
        assert len(e) == (cells == nd).sum()


def test_problem_known_answers_and_views():
    """tests/test_matrix_view.py KATs through the product Problem."""
    import feax_amd as fe
    mesh = fe.StructuredGrid((2, 2, 2), spacing=(5.0, 5.0, 5.0)).to_mesh()
    for view, nI in (("FULL", 6912), ("UPPER", 3600), ("LOWER", 3600)):
        p = LE()(mesh, vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], 10.0, 1e-5)], matrix_view=view)
        assert len(p.I) == len(p.J) == 6912
        assert len(p.I_filtered) == nI
        assert p.num_total_dofs_all_vars == 81 and p.csr_nse == 3087
        indptr, indices, slots = p.view_pattern()
        If, Jf, _ = st.filter_view(p.I, p.J, view)
        cs = st.csr_structure(If, Jf, 81)
        assert np.array_equal(indptr, cs["indptr"]) and np.array_equal(indices, cs["indices"])
    with pytest.raises(ValueError):
        LE()(mesh, vec=3, dim=3, matrix_view="DIAG")


def test_unsupported_problems_raise():
    import feax_amd as fe
    mesh = fe.StructuredGrid((1, 1, 1)).to_mesh()
    with pytest.raises(NotImplementedError):
        fe.Problem(mesh, vec=2, dim=3)
    assert fe.Problem(fe.mesh.rectangle_mesh(1, 1), vec=1, dim=2, ele_type="QUAD4").kernel_tag == "quad4s"
    tets = fe.Mesh(np.vstack([np.zeros(3), np.eye(3)]), np.array([[0, 1, 2, 3]]), ele_type="TET4")
    assert fe.Problem(tets, vec=1, dim=3, ele_type="TET4").kernel_tag == "tet4s"
    with pytest.raises(NotImplementedError):
        fe.Problem(tets, vec=2, dim=3, ele_type="TET4")
    with pytest.raises(ValueError, match=r"\(E, 20\)"):
        fe.Problem(mesh, vec=3, dim=3, ele_type="HEX20")
    with pytest.raises(NotImplementedError):
        fe.Problem(mesh, vec=3, dim=3, ele_type="TET10")
    assert fe.Problem(mesh, vec=1, dim=3).kernel_tag == "hex8s"
    with pytest.raises(ValueError, match=r"\(E, 4\)"):
        fe.Problem(mesh, vec=3, dim=3, ele_type="TET4")


def test_dirichlet_bc_matches_oracle():
    import feax_amd as fe
    pts, cells = st.structured_grid((4, 3, 2), (1.0, 1.0, 1.0))
    p = fe.Problem(fe.Mesh(pts, cells), vec=3, dim=3)
    specs = [fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 0.0),
             fe.DirichletBCSpec(lambda x: np.isclose(x[1], 0.0), "z", lambda x: 0.1 * x[0])]
    bc = fe.DirichletBCConfig(specs).create_bc(p)
    rows, vals = st.dirichlet_rows(pts, [(lambda x: np.isclose(x[0], 0.0), "all", 0.0),
                                         (lambda x: np.isclose(x[1], 0.0), 2, lambda x: 0.1 * x[0])])
    assert np.array_equal(bc.bc_rows, rows) and np.allclose(bc.bc_vals, vals)
    assert bc.bc_mask.sum() == len(rows) and np.all(bc.bc_mask[rows])
    bc2 = bc.replace_vals(bc.bc_vals * 2)
    assert np.array_equal(bc2.bc_rows, bc.bc_rows) and np.allclose(bc2.bc_vals, 2 * vals)
    with pytest.raises(ValueError):
        fe.DirichletBCSpec(lambda x: True, "w", 0.0)
    u = fe.zero_like_initial_guess(p, bc)
    assert np.allclose(u[bc.bc_rows], bc.bc_vals) and np.count_nonzero(u) == np.count_nonzero(bc.bc_vals)


def test_traced_params_creators_and_solution():
    import feax_amd as fe
    mesh = fe.StructuredGrid((3, 2, 2)).to_mesh()
    p = fe.Problem(mesh, vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], 3.0)])
    assert fe.TracedParams.create_node_var(p, 2.0).shape == (p.num_total_nodes,)
    assert fe.TracedParams.create_cell_var(p, 2.0).shape == (p.num_cells,)
    assert fe.TracedParams.create_uniform_surface_var(p, 1.0).shape == (4, 4)
    cv = fe.TracedParams.create_cell_var_from_fn(p, lambda x: x[0])
    assert np.allclose(np.sort(np.unique(cv)), [0.5, 1.5, 2.5])
    tp = fe.TracedParams((np.ones(3),))
    assert tp.surface_vars == []
    sol = fe.Solution(np.arange(3.0 * p.num_total_nodes), ((p.num_total_nodes, 3),))
    assert sol.field(0).shape == (p.num_total_nodes, 3)
    assert np.array_equal(sol.node_var(1), np.arange(1.0, 3.0 * p.num_total_nodes, 3))
    with pytest.raises(ValueError):
        sol.node_var()


def test_face_incidence_and_block3_check():
    import feax_amd as fe
    from feax_amd import _native as nat
    mesh = fe.StructuredGrid((3, 2, 2)).to_mesh()
    p = fe.Problem(mesh, vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], 3.0),
                                                     lambda x: np.isclose(x[2], 0.0)])
    F = len(p.all_faces)
    assert F == 4 + 6
    fi = p.face_inc
    f, a = fi >> 3, fi & 7
    from feax_amd.fe import HEX8_FACE_INDS
    for k in range(len(fi)):
        e, lf = p.all_faces[f[k]]
        assert a[k] in HEX8_FACE_INDS[lf]
    assert len(fi) == 4 * F
    out = ctypes.c_int32(-1)
    nat.check(nat.lib().fx_csr_is_block3(p.num_total_dofs_all_vars, nat.ptr(p.csr_indptr),
                                         nat.ptr(p.csr_indices), ctypes.addressof(out)))
    assert out.value == 1
    bad = p.csr_indices.copy()
    bad[[1, 2]] = bad[[2, 1]]
    nat.check(nat.lib().fx_csr_is_block3(p.num_total_dofs_all_vars, nat.ptr(p.csr_indptr), nat.ptr(bad),
                                         ctypes.addressof(out)))
    assert out.value == 0


def test_native_errors_are_reported():
    from feax_amd import _native as nat
    L = nat.lib()
    nse = ctypes.c_int64(0)
    cells = np.array([[0, 1, 2, 3, 4, 5, 6, 99]], np.int32)
    indptr = np.zeros(3 * 8 + 1, np.int32)
    rc = L.fx_hex8_pattern(8, 1, nat.ptr(cells), ctypes.addressof(nse), None, nat.ptr(indptr), None, None, None, None)
    assert rc == 1001
    assert "out of range" in nat.last_error()
    with pytest.raises(nat.NativeError):
        nat.check(rc, "fx_hex8_pattern")


def test_transpose_maps_and_volume_slots():
    """csr_T_* reproduce scipy's transpose of a random-valued matrix on the
    pattern; csr_volume_slots equals the oracle's slot map (feax's
    csr_full_slot volume block)."""
    import scipy.sparse as sp
    from conftest import cantilever
    from oracle import hex8
    p, _, _ = cantilever((4, 3, 2), (1.0, 1.0, 1.0), 4.0)
    n = p.num_total_dofs_all_vars
    rng = np.random.default_rng(0)
    data = rng.standard_normal(p.csr_nse)
    A = sp.csr_matrix((data, p.csr_indices, p.csr_indptr), shape=(n, n))
    At = A.T.tocsr()
    At.sort_indices()
    np.testing.assert_array_equal(p.csr_T_indptr, At.indptr)
    np.testing.assert_array_equal(p.csr_T_indices, At.indices)
    np.testing.assert_array_equal(data[p.csr_T_perm], At.data)
    m = hex8.Model(p.points, p.cells.astype(np.int64), p.boundary_inds_list)
    np.testing.assert_array_equal(p.csr_volume_slots, m.csr["volume_slots"])


def test_structured_dims_detects_box_grids_only():
    import feax_amd as fe
    from feax_amd.solvers.gmg import coarse_dims, structured_dims
    prob = fe.Problem(fe.StructuredGrid((6, 4, 5), spacing=(0.5, 1.0, 2.0)).to_mesh(), vec=3, dim=3)
    dims, axes = structured_dims(prob)
    assert dims == (6, 4, 5) and np.allclose(axes[2], 2.0 * np.arange(6))
    assert coarse_dims((6, 4, 5), 2) == (3, 2, 5)          # even counts first
    assert coarse_dims((3, 2, 5), 2) == (2, 2, 3)          # then odd ones, ceil-halved
    assert coarse_dims((2, 2, 2), 2) is None
    pts = prob.points[::-1].copy()
    cells = len(pts) - 1 - prob.cells
    prob2 = fe.Problem(fe.Mesh(pts, cells), vec=3, dim=3)
    with pytest.raises(NotImplementedError, match="structured box"):
        structured_dims(prob2)


def test_dirichlet_values_may_be_a_tensor():
    """replace_vals keeps a torch tensor (so a solve can differentiate with
    respect to it); vals_numpy and zero_like_initial_guess read it; a wrong
    length is refused."""
    import torch
    import feax_amd as fe
    mesh = fe.mesh.rectangle_mesh(2, 2)

    class P(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.25, E=2.5)

    prob = P(mesh, vec=2, dim=2, ele_type="QUAD4")
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    t = torch.linspace(0.1, 0.6, len(bc.bc_rows), dtype=torch.float64, requires_grad=True)
    bc2 = bc.replace_vals(t)
    assert bc2.bc_vals is t and np.allclose(bc2.vals_numpy(), t.detach().numpy())
    u = fe.zero_like_initial_guess(prob, bc2)
    assert np.allclose(u[bc.bc_rows], t.detach().numpy()) and np.count_nonzero(u) == len(bc.bc_rows)
    with pytest.raises(ValueError):
        bc.replace_vals(torch.zeros(len(bc.bc_rows) + 1))


def test_solution_behaves_as_the_flat_dof_vector():
    """Arithmetic, indexing and torch functions see the flat DOFs
    (feax/solution.py:60-170)."""
    import torch
    from feax_amd.solution import Solution
    d = torch.arange(12, dtype=torch.float64)
    s = Solution(d, ((4, 3),))
    assert torch.equal(s ** 2, d ** 2) and torch.equal(s / 2, d / 2) and torch.equal(2 / (s + 1), 2 / (d + 1))
    assert torch.equal(s - s, torch.zeros(12, dtype=torch.float64)) and float(s @ s) == float(d @ d)
    assert float(torch.sum(s)) == 66.0 and float(torch.dot(s, d)) == float(d @ d)
    assert torch.equal(torch.stack([s, d]), torch.stack([d, d]))
    assert s.field(0).shape == (4, 3) and torch.equal(s.node_var(1), d.reshape(4, 3)[:, 1])
    assert np.array_equal(np.asarray(s), d.numpy()) and s.block_until_ready() is s


def test_jacobian_info_without_assembly():
    import feax_amd as fe
    mesh = fe.StructuredGrid((3, 2, 2)).to_mesh()
    p = fe.Problem(mesh, vec=3, dim=3)
    info = fe.get_jacobian_info(p, [np.zeros((p.num_total_nodes, 3))], None)
    assert info["shape"] == (p.num_total_dofs_all_vars,) * 2 and info["nnz"] == p.csr_nse
    assert info["matrix_view"] == fe.MatrixView.FULL
    pu = fe.Problem(mesh, vec=3, dim=3, matrix_view=fe.MatrixView.UPPER)
    n = pu.num_total_dofs_all_vars
    assert fe.get_jacobian_info(pu, None, None)["nnz"] == (p.csr_nse + n) // 2
