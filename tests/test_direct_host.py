"""Host side of the direct solver: the band ordering and the options
(DirectSolverOptions, feax/solvers/options.py:595-635).  CPU only."""

import numpy as np
import pytest

from test_structure_host import shuffled_mesh


class _LE:
    @staticmethod
    def problem(pts, cells):
        import feax_amd as fe

        class LE(fe.Problem):
            def get_tensor_map(self):
                return fe.materials.IsotropicLinearElastic(nu=0.3)

        return LE(fe.Mesh(pts, cells), vec=3, dim=3)


@pytest.mark.parametrize("dims", [(4, 2, 3), (12, 8, 8), (50, 5, 5)])
def test_bandwidth_of_the_grid_numbering_is_the_closed_form(dims):
    """StructuredGrid numbers nodes (i (ny+1) + j)(nz+1) + k: the farthest
    coupled DOFs are a node's neighbour one step in all three directions,
    3 ((ny+1)(nz+1) + (nz+1) + 1) + 2 DOFs apart."""
    import feax_amd as fe
    from feax_amd.solvers.direct import band_ordering
    mesh = fe.StructuredGrid(dims).to_mesh()
    p = _LE.problem(mesh.points, mesh.cells)
    nx, ny, nz = dims
    perm, inv, bw = band_ordering(p)
    closed = 3 * ((ny + 1) * (nz + 1) + (nz + 1) + 1) + 2
    if perm is None:
        assert bw == closed
    else:
        assert bw < closed
        assert np.array_equal(np.sort(perm), np.arange(p.num_total_dofs_all_vars))
        assert np.array_equal(inv[perm], np.arange(p.num_total_dofs_all_vars))


def test_rcm_narrows_the_band_of_a_shuffled_mesh():
    from feax_amd.solvers.direct import band_ordering
    pts, cells = shuffled_mesh((6, 3, 3), 5)
    p = _LE.problem(pts, cells)
    perm, inv, bw = band_ordering(p)
    indptr, indices = np.asarray(p.csr_indptr), np.asarray(p.csr_indices)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    natural = int(np.abs(rows - indices).max())
    assert perm is not None and bw < natural / 2
    # node-wise: the 3 DOFs of a node stay consecutive
    assert np.array_equal(perm.reshape(-1, 3) % 3, np.tile(np.arange(3), (len(perm) // 3, 1)))
    assert bw == int(np.abs(inv[rows].astype(np.int64) - inv[indices]).max())


def test_direct_options_validate_and_route():
    import feax_amd as fe
    assert not fe.DirectSolverOptions().uses_x0()
    with pytest.raises(ValueError, match="Invalid direct solver"):
        fe.DirectSolverOptions(solver="mumps")
    for name in ("auto", "cudss", "spsolve", "umfpack", "cholmod"):
        fe.DirectSolverOptions(solver=name)
    mesh = fe.StructuredGrid((2, 2, 2)).to_mesh()
    p = _LE.problem(mesh.points, mesh.cells)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 0.0)]).create_bc(p)
    with pytest.raises(NotImplementedError, match="symmetric"):
        fe.create_solver(p, bc, fe.DirectSolverOptions(solver="cudss"), linear=False, symmetric_elimination=False)
    # feax/solver.py:316-321: "auto" (and missing options) need a sample traced_params
    for opts in (None, fe.DirectSolverOptions(), fe.KrylovSolverOptions()):
        with pytest.raises(ValueError, match="traced_params is required"):
            fe.create_solver(p, bc, opts, linear=True)
    with pytest.raises(ValueError, match="traced_params is required"):
        fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg"), fe.KrylovSolverOptions(), linear=True)
    with pytest.raises(ValueError, match="unresolved"):
        fe.create_direct_solve_fn(fe.DirectSolverOptions())
    with pytest.raises(NotImplementedError, match="GENERAL"):
        fe.resolve_direct_solver(fe.DirectSolverOptions(), fe.MatrixProperty.GENERAL)
    assert fe.resolve_direct_solver(fe.DirectSolverOptions(), fe.MatrixProperty.SPD).solver == "cudss"
    with pytest.raises(TypeError, match="Expected"):
        fe.create_solver(p, bc, object(), linear=True)
