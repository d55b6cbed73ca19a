"""Oracle: mesh numbering, boundary faces and the CSR assembly structure.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Integer-exact restatement of:
  feax/spgrid.py:21-25, 109-170      StructuredGrid numbering, to_mesh
  feax/basis.py:64-70, 244-325       HEX8 facet vertex lists (basix facets,
                                     meshio re-order [0,1,3,2,4,5,7,6])
  feax/fe.py:267-318                 boundary faces = argwhere(all 4 vertices
                                     satisfy location_fn) in (cell, face) order
  feax/problem.py:161-231            DOF map, volume-then-face COO (I, J),
                                     UPPER/LOWER filter
  feax/problem.py:316-484            argsort -> dedupe -> indptr / indices,
                                     slot maps, diag slots, transpose maps
  feax/DCboundary.py:395-501         Dirichlet rows (sorted, unique)

Pinned by: the reference's known answers (len(I)=6912, UPPER/LOWER 3600,
shape (81, 81) for the 2x2x2 fixture, tests/test_matrix_view.py:107-133,
174-175, 199-200), the closed form nse = 9*prod(3(n_d+1)-2) for structured
boxes, len(I) = 576*(E+F), and a brute-force dense adjacency check on tiny
meshes (tests/test_oracle.py, tests/test_structure_host.py); location and
value functions are evaluated point by point, and pinned against
explicit per-point loops (tests/test_structure_host.py).
"""

import numpy as np

# feax/spgrid.py:21-25 -- local HEX8 corner offsets (dx, dy, dz) from the
# cell's base corner, in feax's HEX8 (gmsh/meshio) local order.
HEX8_OFFSETS = np.array([(0, 1, 1), (0, 0, 1), (0, 0, 0), (0, 1, 0),
                         (1, 1, 1), (1, 0, 1), (1, 0, 0), (1, 1, 0)], dtype=np.int64)

# basix hexahedron facets (vertex numbering of the basix reference cell,
# x fastest) -> feax local numbering through the involutive re-order
# [0,1,3,2,4,5,7,6] (feax/basis.py:65-66, 321-322).
_BASIX_HEX_FACETS = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6),
                     (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]
_REORDER = np.array([0, 1, 3, 2, 4, 5, 7, 6])
HEX8_FACE_INDS = np.array([[int(np.argwhere(_REORDER == v)[0, 0]) for v in f]
                           for f in _BASIX_HEX_FACETS], dtype=np.int64)


def structured_grid(dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """Return (points (N,3) f64, cells (E,8) int64) of a full StructuredGrid.

    feax/spgrid.py:109-130 (cell_ijk, cell_to_nodes, node_coords) and
    156-170 (to_mesh: unique + searchsorted renumbering, the identity for
    the full grid).
    """
    nx, ny, nz = (int(d) for d in dims)
    spacing = np.asarray(spacing, float)
    origin = np.asarray(origin, float)
    e = np.arange(nx * ny * nz, dtype=np.int64)
    nynz = ny * nz
    cx, r = e // nynz, e % nynz
    cy, cz = r // nz, r % nz
    ny1, nz1 = ny + 1, nz + 1
    I = cx[:, None] + HEX8_OFFSETS[None, :, 0]
    J = cy[:, None] + HEX8_OFFSETS[None, :, 1]
    K = cz[:, None] + HEX8_OFFSETS[None, :, 2]
    cells = (I * ny1 + J) * nz1 + K
    used = np.unique(cells)
    n = used
    i, rr = n // (ny1 * nz1), n % (ny1 * nz1)
    j, k = rr // nz1, rr % nz1
    points = origin + np.stack([i, j, k], axis=-1) * spacing
    return points, np.searchsorted(used, cells)


def _location_per_node(points, location_fn):
    """Boolean flag of every node, ``location_fn`` called once per point as
    jax.vmap calls it (feax/fe.py:288-301, feax/DCboundary.py:448-460): a
    one-argument function gets the (dim,) coordinates, a two-argument one
    also the node index."""
    nargs = location_fn.__code__.co_argcount if hasattr(location_fn, "__code__") else 1
    if nargs == 1:
        return np.array([bool(location_fn(p)) for p in points], dtype=bool)
    if nargs == 2:
        return np.array([bool(location_fn(p, i)) for i, p in enumerate(points)], dtype=bool)
    raise ValueError(f"Wrong number of arguments for location_fn: must be 1 or 2, got {nargs}")


def boundary_faces(points, cells, location_fn):
    """(F, 2) int64 array of (cell, local face) whose 4 vertices all satisfy
    ``location_fn`` -- feax/fe.py:267-318 (np.argwhere => (cell, face) order).

    ``location_fn`` is evaluated point by point.  A face vertex is a mesh
    node, so evaluating each node once and gathering the flags per face
    vertex gives what feax's vmap over (cell, face, vertex) gives.
    """
    flag = _location_per_node(points, location_fn)
    return np.argwhere(flag[cells[:, HEX8_FACE_INDS]].all(axis=2)).astype(np.int64)


def build_inds(cells, vec=3, offset=0):
    """Per-element global DOF indices (E, 8*vec) int32 -- problem.py:161-181."""
    c = np.asarray(cells)
    dof = vec * c[:, :, None] + np.arange(vec)[None, None, :] + offset
    return dof.reshape(len(c), -1).astype(np.int32)


def coo_ij(cells, boundary_inds_list, vec=3):
    """Raw COO (I, J) in volume-then-face order + residual scatter DOFs.

    feax/problem.py:183-212: each element contributes repeat(inds) rows x cols
    (row-major r*ndof+c); each selected face contributes its cell's full DOF
    block; boundaries are appended in location_fns order.
    """
    inds = build_inds(cells, vec)
    nd = inds.shape[1]
    I = [np.repeat(inds[:, :, None], nd, axis=2).reshape(-1)]
    J = [np.repeat(inds[:, None, :], nd, axis=1).reshape(-1)]
    res = [inds.reshape(-1)]
    for bi in boundary_inds_list:
        fi = build_inds(np.asarray(cells)[bi[:, 0]], vec)
        I.append(np.repeat(fi[:, :, None], nd, axis=2).reshape(-1))
        J.append(np.repeat(fi[:, None, :], nd, axis=1).reshape(-1))
        res.append(fi.reshape(-1))
    return np.concatenate(I), np.concatenate(J), np.concatenate(res)


def filter_view(I, J, view="FULL"):
    """UPPER keeps j >= i, LOWER keeps j <= i (feax/problem.py:218-231)."""
    if view == "UPPER":
        m = J >= I
    elif view == "LOWER":
        m = J <= I
    else:
        return I, J, None
    return I[m], J[m], np.where(m)[0].astype(np.int32)


def csr_structure(I, J, n, I_all=None, J_all=None, num_cells=None, ndof_vol=24):
    """The argsort-based CSR structure of feax/problem.py:316-484.

    ``I, J`` are the (filtered) raw entries; ``I_all, J_all`` the unfiltered
    ones (default: the same).  Returns a dict with indptr, indices,
    row_of_slot, slot (raw -> slot), diag_slots, T_perm, T_indptr,
    T_indices, structurally_symmetric, full_slot (unfiltered -> slot | nse)
    and volume_slots (E, ndof^2).
    """
    rows = np.asarray(I).astype(np.int64)
    cols = np.asarray(J).astype(np.int64)
    keys = rows * n + cols
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    is_new = np.ones(len(sk), dtype=bool)
    if len(sk) > 1:
        is_new[1:] = sk[1:] != sk[:-1]
    seg = np.cumsum(is_new) - 1
    nse = int(seg[-1]) + 1 if len(sk) else 0
    uk = sk[is_new]
    csr_rows = (uk // n).astype(np.int32)
    csr_cols = (uk % n).astype(np.int32)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, csr_rows.astype(np.int64) + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    slot = np.empty(len(sk), dtype=np.int64)
    slot[order] = seg
    diag = np.full(n, -1, dtype=np.int64)
    dpos = np.where(csr_rows == csr_cols)[0]
    diag[csr_rows[dpos]] = dpos
    if I_all is None:
        I_all, J_all = I, J
    ka = np.asarray(I_all).astype(np.int64) * n + np.asarray(J_all).astype(np.int64)
    pos = np.searchsorted(uk, ka)
    pos_safe = np.minimum(pos, max(nse - 1, 0))
    full_slot = np.where(uk[pos_safe] == ka, pos, nse).astype(np.int64)
    tk = csr_cols.astype(np.int64) * n + csr_rows.astype(np.int64)
    t_order = np.argsort(tk, kind="stable")
    t_rows = csr_cols[t_order]
    t_cols = csr_rows[t_order]
    t_indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(t_indptr, t_rows.astype(np.int64) + 1, 1)
    t_indptr = np.cumsum(t_indptr).astype(np.int32)
    out = dict(
        nse=nse, indptr=indptr, indices=csr_cols, row_of_slot=csr_rows,
        slot=slot.astype(np.int32), diag_slots=diag.astype(np.int32),
        T_perm=t_order.astype(np.int32), T_indptr=t_indptr,
        T_indices=t_cols.astype(np.int32),
        structurally_symmetric=bool(np.array_equal(csr_cols, t_cols)
                                    and np.array_equal(indptr, t_indptr)),
        full_slot=full_slot.astype(np.int32))
    if num_cells is not None:
        vl = num_cells * ndof_vol * ndof_vol
        out["volume_slots"] = out["full_slot"][:vl].reshape(num_cells, ndof_vol * ndof_vol)
    return out


def dirichlet_rows(points, specs, vec=3):
    """(bc_rows int32 sorted unique, bc_vals f64) -- DCboundary.py:395-501.

    ``specs`` is a list of (location_fn, component, value) where component is
    an int or 'all' and value a float or a fn of one point.  Location and
    value functions are called point by point (the jax.vmap of
    feax/DCboundary.py:457-470).  Rows are concatenated per spec and
    component, stably sorted, and duplicates keep their first occurrence
    (np.unique return_index).
    """
    rows, vals = [], []
    for loc, comp, value in specs:
        comps = list(range(vec)) if comp == "all" else [int(comp)]
        for c in comps:
            nodes = np.argwhere(_location_per_node(points, loc)).reshape(-1)
            if len(nodes) == 0:
                continue
            rows.append(nodes * vec + c)
            if callable(value):
                vals.append(np.array([float(np.asarray(value(p)).reshape(())) for p in points[nodes]]))
            else:
                vals.append(np.full(len(nodes), float(value)))
    if not rows:
        return np.zeros(0, np.int32), np.zeros(0, np.float64)
    r = np.concatenate(rows)
    v = np.concatenate(vals)
    s = np.argsort(r, kind="stable")      # jnp.argsort is stable by default
    r, v = r[s], v[s]
    ur, ui = np.unique(r, return_index=True)
    return ur.astype(np.int32), v[ui]


def structured_nse(dims):
    """Closed-form CSR nnz of a full structured HEX8 box, vec=3."""
    p = 1
    for d in dims:
        p *= 3 * (int(d) + 1) - 2
    return 9 * p
