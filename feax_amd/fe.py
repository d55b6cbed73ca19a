"""HEX8 / TET4 / QUAD4 reference-element facts needed on the host.

Geometry itself (shape gradients, JxW, Nanson scale) is never stored: the
HIP kernels recompute it per Gauss point from the node coordinates.  The
host only needs the facet vertex lists to find the faces on a boundary
(feax/fe.py:267-318).
"""

import numpy as np

# Local vertices of the 6 HEX8 facets in feax numbering: basix hexahedron
# facets through the re-order [0,1,3,2,4,5,7,6] (feax/basis.py:65-66, 321-322).
HEX8_FACE_INDS = np.array([[0, 1, 3, 2], [0, 1, 4, 5], [0, 3, 4, 7],
                           [1, 2, 5, 6], [3, 2, 7, 6], [4, 5, 7, 6]], dtype=np.int64)
NUM_QUADS = 8        # 2x2x2 Gauss points
NUM_FACE_QUADS = 4   # 2x2 per facet

# TET4 (feax/basis.py:78-83): basix tetrahedron facets, facet i opposite
# vertex i; one Gauss point per cell and per facet.
TET4_FACE_INDS = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]], dtype=np.int64)
# QUAD4 (feax/basis.py:90-95): basix quadrilateral edges through the
# re-order [0, 1, 3, 2]; 2x2 Gauss points, 2 per edge; dim = vec = 2.
QUAD4_FACE_INDS = np.array([[0, 1], [0, 3], [1, 2], [3, 2]], dtype=np.int64)
ELEMENTS = {"HEX8": dict(npe=8, faces=HEX8_FACE_INDS, num_quads=NUM_QUADS, num_face_quads=NUM_FACE_QUADS, dim=3),
            "TET4": dict(npe=4, faces=TET4_FACE_INDS, num_quads=1, num_face_quads=1, dim=3),
            "QUAD4": dict(npe=4, faces=QUAD4_FACE_INDS, num_quads=4, num_face_quads=2, dim=2)}


def element_info(ele_type):
    if ele_type not in ELEMENTS:
        raise NotImplementedError(f"the HIP path covers {sorted(ELEMENTS)} (got {ele_type})")
    return ELEMENTS[ele_type]


def eval_points(fn, pts):
    """Evaluate a point predicate / function on many points.

    feax vmaps user location functions over points (feax/fe.py:297-301).
    Here ``fn`` is first called once on the (dim, M) coordinate array (so
    ``p[0]`` is the x column and numpy ufuncs vectorise); if that does not
    give M values it is called point by point.
    """
    pts = np.asarray(pts, float)
    if len(pts) == 0:
        return np.zeros(0)
    try:
        out = np.asarray(fn(pts.T))
        if out.shape == (len(pts),):
            return out
        if out.ndim == 0:                 # constant function
            return np.full(len(pts), out[()])
    except Exception:
        pass
    return np.asarray([fn(p) for p in pts])


def boundary_conditions_inds(points, cells, location_fns, ele_type="HEX8"):
    """List of (F, 2) int arrays (cell, local face) whose vertices all satisfy
    each location_fn, in np.argwhere (cell, face) order -- feax/fe.py:267-318.
    Two-argument location functions receive (point, node index)."""
    out = []
    if not location_fns:
        return out
    fverts = cells[:, element_info(ele_type)["faces"]]     # (E, faces, verts) node ids
    for fn in location_fns:
        nargs = fn.__code__.co_argcount if hasattr(fn, "__code__") else 1
        nodes = np.arange(len(points))
        if nargs == 1:
            flag = eval_points(fn, points).astype(bool)
        elif nargs == 2:
            flag = np.asarray([bool(fn(p, i)) for p, i in zip(points, nodes)])
        else:
            raise ValueError(f"Wrong number of arguments for location_fn: must be 1 or 2, get {nargs}")
        on = flag[fverts].all(axis=2)
        out.append(np.argwhere(on).astype(np.int64))
    return out
