"""Host-side structure of batched instances (no GPU): the concatenated
problem's CSR pattern is the instances' patterns placed block-diagonally,
loaded faces and Dirichlet rows carry the node/cell offsets."""

import numpy as np

from conftest import cantilever


def _instances():
    dims = [((6, 3, 3), 6.0), ((4, 2, 5), 4.0), ((3, 4, 2), 3.0)]
    return [cantilever(d, (1.0, 1.0, 1.0), L, material="cell") for d, L in dims]


def test_batched_pattern_is_block_diagonal():
    from feax_amd.batch import BatchedProblem
    inst = _instances()
    probs = [p for p, _, _ in inst]
    bp = BatchedProblem(probs)
    assert bp.nseg == 3
    assert list(bp.seg_node_ptr) == list(np.cumsum([0] + [p.num_total_nodes for p in probs]))
    ip, ix = [0], []
    for p, off in zip(probs, bp.seg_node_ptr[:-1]):
        ix.append(p.csr_indices + 3 * off)
        ip.extend((p.csr_indptr[1:] + ip[-1]).tolist())
    np.testing.assert_array_equal(bp.csr_indptr, np.array(ip))
    np.testing.assert_array_equal(bp.csr_indices, np.concatenate(ix))
    assert bp.csr_nse == sum(p.csr_nse for p in probs)
    # loaded faces: the instances' faces with cell offsets
    faces = np.concatenate([p.boundary_inds_list[0] + np.array([c, 0]) for p, c in zip(probs, bp.seg_cell_ptr[:-1])])
    np.testing.assert_array_equal(bp.boundary_inds_list[0], faces)


def test_batched_bc_and_params():
    from feax_amd.batch import BatchedProblem, batch_bc, batch_params
    inst = _instances()
    bp = BatchedProblem([p for p, _, _ in inst])
    bc = batch_bc(bp, [b for _, b, _ in inst])
    rows = np.concatenate([b.bc_rows + 3 * o for (_, b, _), o in zip(inst, bp.seg_node_ptr[:-1])])
    np.testing.assert_array_equal(bc.bc_rows, rows)
    assert np.all(np.diff(bc.bc_rows) > 0)
    assert bc.bc_mask.sum() == len(rows) and bc.total_dofs == bp.num_total_dofs_all_vars
    tp = batch_params(bp, [t for _, _, t in inst])
    assert tp.volume_vars[0].shape == (bp.num_cells,)
    assert tp.surface_vars[0][0].shape == (len(bp.boundary_inds_list[0]), 4)


def test_batched_rejects_mixed_laws():
    import pytest
    import feax_amd as fe
    from feax_amd.batch import BatchedProblem
    p1, _, _ = cantilever((3, 2, 2), (1.0, 1.0, 1.0), 3.0)

    class Other(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.25)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    p2 = Other(fe.StructuredGrid((3, 2, 2)).to_mesh(), vec=3, dim=3, location_fns=[lambda p: np.isclose(p[0], 3.0)])
    with pytest.raises(ValueError):
        BatchedProblem([p1, p2])
