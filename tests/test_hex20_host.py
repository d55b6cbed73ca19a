"""HEX20 host structure: the structured mesh builder against the oracle's,
the native CSR pattern (fx_hex20_pattern) bit-equal to the oracle's argsort
structure over the 3600-entry element blocks, the element ranks, the face
selection by corners and the face incidence over each facet's 8 nodes."""

import numpy as np

from oracle import hex20, structure as st


def _problem(dims, **kw):
    import feax_amd as fe
    mesh = fe.hex20_mesh(dims, spacing=(1.0, 0.5, 2.0))
    return fe.Problem(mesh, vec=3, dim=3, ele_type="HEX20", **kw)


def test_mesh_matches_oracle_box():
    import feax_amd as fe
    m = fe.hex20_mesh((3, 2, 4), spacing=(1.0, 0.5, 2.0))
    pts, cells = hex20.box((3, 2, 4), (1.0, 0.5, 2.0))
    assert np.array_equal(m.cells, cells) and np.array_equal(m.points, pts)


def test_pattern_ranks_and_faces_match_oracle():
    p = _problem((3, 2, 2), location_fns=[lambda x: np.isclose(x[0], 3.0), lambda x: np.isclose(x[2], 0.0)])
    cells = p.cells.astype(np.int64)
    I, J, _ = st.coo_ij(cells, p.boundary_inds_list)
    csr = st.csr_structure(I, J, p.num_total_dofs_all_vars, num_cells=p.num_cells, ndof_vol=60)
    assert p.csr_nse == csr["nse"]
    assert np.array_equal(p.csr_indptr, csr["indptr"]) and np.array_equal(p.csr_indices, csr["indices"])
    # rank of node b in node a's neighbour list <-> the oracle's volume slots
    vs = csr["volume_slots"].reshape(p.num_cells, 20, 3, 20, 3)
    rank = p.elem_rank.reshape(p.num_cells, 20, 20).astype(np.int64)
    row0 = p.csr_indptr[3 * cells]                                    # (E, 20)
    expect = row0[:, :, None] + 3 * rank                               # slot of (a, i=0; b, j=0)
    assert np.array_equal(vs[:, :, 0, :, 0], expect)
    # node incidence: (e << 5) | a, ascending elements per node
    inc = p.node_inc
    for nd in (0, 5, p.num_total_nodes - 1):
        ent = inc[p.node_inc_ptr[nd]:p.node_inc_ptr[nd + 1]]
        e, a = ent >> 5, ent & 31
        assert np.all(np.diff(e) > 0) and np.all(cells[e, a] == nd)
    # faces: by corners, as the oracle; face incidence lists each face's 8 nodes
    ob = [hex20.boundary_faces(p.points, cells, fn) for fn in p.location_fns]
    for a, b in zip(p.boundary_inds_list, ob):
        assert np.array_equal(a, b)
    cnt = np.zeros(p.num_total_nodes, np.int64)
    for f, (e, lf) in enumerate(p.all_faces):
        cnt[cells[e, hex20.FACE_NODES[lf]]] += 1
    assert np.array_equal(np.diff(p.face_inc_ptr), cnt)
    fi = p.face_inc
    f_, a_ = fi >> 5, fi & 31
    owners = np.repeat(np.arange(p.num_total_nodes), np.diff(p.face_inc_ptr))
    assert np.all(cells[p.all_faces[f_, 0], a_] == owners)
    assert p.csr_max_degree == 81 or p.csr_max_degree <= 96
