"""HEX8 / HEX20 / TET4 / QUAD4 reference-element facts needed on the host.

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
# HEX20 (feax/basis.py:71-77): basix serendipity degree 2 in meshio order
# (8 corners, then the edge nodes of these corner pairs); boundary faces are
# selected by their 4 corners, as for HEX8; 2x2x2 Gauss points.
HEX20_EDGES = np.array([(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                        (0, 4), (1, 5), (2, 6), (3, 7)], dtype=np.int64)
ELEMENTS = {"HEX8": dict(npe=8, faces=HEX8_FACE_INDS, num_quads=NUM_QUADS, num_face_quads=NUM_FACE_QUADS, dim=3),
            "HEX20": dict(npe=20, faces=HEX8_FACE_INDS, num_quads=NUM_QUADS, num_face_quads=NUM_FACE_QUADS, dim=3),
            "TET4": dict(npe=4, faces=TET4_FACE_INDS, num_quads=1, num_face_quads=1, dim=3),
            "QUAD4": dict(npe=4, faces=QUAD4_FACE_INDS, num_quads=4, num_face_quads=2, dim=2)}


_G2 = ((1.0 - 1.0 / np.sqrt(3.0)) / 2.0, (1.0 + 1.0 / np.sqrt(3.0)) / 2.0)   # 2-point Gauss on [0, 1]
# Reference-cell vertices in feax local order.
_REF_NODES = {
    "HEX8": np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                      (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], float),
    "TET4": np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], float),
    "QUAD4": np.array([(0, 0), (1, 0), (1, 1), (0, 1)], float),
}


def _hex8_face_ref(lf, s, t):
    return [(s, t, 0.0), (s, 0.0, t), (0.0, s, t), (1.0, s, t), (s, 1.0, t), (s, t, 1.0)][lf]


def _hex20_values(xi):
    """HEX20 basis values (M, 20): the 20-node nodal serendipity functions
    psi re-combined to basix's DOFs (corner values, edge means):
    phi_edge = 3/2 psi_edge, phi_corner = psi_corner - 1/4 sum of its edges' psi."""
    V = _REF_NODES["HEX8"]
    s = 2.0 * np.atleast_2d(xi) - 1.0
    sg = 2.0 * V - 1.0
    corner = np.prod(1.0 + sg[None] * s[:, None], axis=2) * ((sg[None] * s[:, None]).sum(axis=2) - 2.0) / 8.0
    edge = np.empty((len(s), 12))
    for k, (p, q) in enumerate(HEX20_EDGES):
        ax = int(np.flatnonzero(V[p] != V[q])[0])
        i, j = [d for d in range(3) if d != ax]
        edge[:, k] = 0.25 * (1.0 - s[:, ax] ** 2) * (1.0 + sg[p, i] * s[:, i]) * (1.0 + sg[p, j] * s[:, j])
    phi = np.concatenate([corner, 1.5 * edge], axis=1)
    for v in range(8):
        at = [k for k, (p, q) in enumerate(HEX20_EDGES) if v in (p, q)]
        phi[:, v] -= 0.25 * edge[:, at].sum(axis=1)
    return phi


def reference_quad_points(ele_type):
    """Volume quadrature points (nq, dim) on the reference cell in the order
    the kernels use for quad variables (basix's: the first coordinate
    slowest; feax/basis.py:155-197)."""
    if ele_type in ("HEX8", "HEX20"):
        return np.array([(a, b, c) for a in _G2 for b in _G2 for c in _G2])
    if ele_type == "TET4":
        return np.array([(0.25, 0.25, 0.25)])
    if ele_type == "QUAD4":
        return np.array([(a, b) for a in _G2 for b in _G2])
    raise NotImplementedError(ele_type)


def reference_face_quad_points(ele_type):
    """Face quadrature points (num_faces, nfq, dim) on the reference cell
    (feax/basis.py:244-325): 2x2 Gauss points on each HEX8 facet, the
    centroid of each TET4 facet, 2 Gauss points along each QUAD4 edge."""
    if ele_type in ("HEX8", "HEX20"):
        return np.array([[_hex8_face_ref(lf, s, t) for s in _G2 for t in _G2] for lf in range(6)])
    V = _REF_NODES[ele_type]
    if ele_type == "TET4":
        return np.array([[V[f].mean(axis=0)] for f in TET4_FACE_INDS])
    if ele_type == "QUAD4":
        return np.array([[V[f[0]] + g * (V[f[1]] - V[f[0]]) for g in _G2] for f in QUAD4_FACE_INDS])
    raise NotImplementedError(ele_type)


def shape_values(ele_type, xi):
    """Shape function values (M, npe) at reference points xi (M, dim)."""
    xi = np.atleast_2d(np.asarray(xi, float))
    if ele_type == "HEX20":
        return _hex20_values(xi)
    if ele_type == "TET4":
        return np.concatenate([1.0 - xi.sum(axis=1, keepdims=True), xi], axis=1)
    V = _REF_NODES[ele_type]
    f = np.where(V[None, :, :] == 1, xi[:, None, :], 1.0 - xi[:, None, :])
    return f.prod(axis=2)


def physical_quad_points(points, cells, ele_type):
    """(E, nq, dim) physical coordinates of the volume quadrature points
    (feax/fe.py:224-239, FiniteElement.get_physical_quad_points)."""
    N = shape_values(ele_type, reference_quad_points(ele_type))
    return np.einsum("qa,ead->eqd", N, np.asarray(points, float)[np.asarray(cells)])


def physical_surface_quad_points(points, cells, boundary_inds, ele_type):
    """(F, nfq, dim) physical coordinates of the face quadrature points of
    the faces ``boundary_inds`` (F, 2) (feax/fe.py:241-265)."""
    ref = reference_face_quad_points(ele_type)
    nf, nfq, dim = ref.shape
    N = shape_values(ele_type, ref.reshape(-1, dim)).reshape(nf, nfq, -1)
    bi = np.asarray(boundary_inds, np.int64).reshape(-1, 2)
    X = np.asarray(points, float)[np.asarray(cells)[bi[:, 0]]]
    return np.einsum("fqa,fad->fqd", N[bi[:, 1]], X)


def element_info(ele_type):
    if ele_type not in ELEMENTS:
        raise NotImplementedError(f"the HIP path covers {sorted(ELEMENTS)} (got {ele_type})")
    return ELEMENTS[ele_type]


_SAMPLE = 48          # points checked against per-point calls before a vectorised result is trusted


def _same(a, b):
    """Exact equality for bool / integer results, 1e-14 relative for floats
    (a vectorised transcendental may differ from its scalar call in the last
    bit); NaN equals NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind in "biu" and b.dtype.kind in "biu":
        return bool(np.array_equal(a, b))
    try:
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    except (TypeError, ValueError):
        return False
    nan = np.isnan(a) & np.isnan(b)
    close = np.abs(a - b) <= 1e-14 * np.maximum(np.abs(a), np.abs(b))
    return bool(np.all(close | nan))


def _per_point(fn, pts, extra):
    return np.asarray([fn(p, *(e[i] for e in extra)) for i, p in enumerate(pts)])


def eval_points(fn, pts, *extra):
    """``fn`` evaluated at every point, as ``jax.vmap(fn)`` does in feax
    (feax/fe.py:297-301, feax/DCboundary.py:457-470,
    feax/traced_params.py:234-375).

    ``pts`` is (M, dim); ``extra`` are per-point arrays passed as further
    arguments (the node index of a two-argument location function).  The
    result has M entries along its first axis.

    Fast path: ``fn`` is called once on the transposed (dim, M) array, so
    that ``p[0]`` is the x column and elementwise numpy code vectorises.  Its
    result is used only if all of these hold:

    * it has exactly M entries (a single value is never copied to all
      points);
    * it equals per-point calls ``fn(p)`` on a seeded sample of the points
      (both outcomes of a predicate included);
    * it equals a call on that sample alone, and its first M entries do not
      change when far-away points are appended to the array (a function that
      reduces over its argument, such as ``np.linalg.norm(p)``, ``np.all``,
      ``np.dot`` or ``np.allclose``, depends on the other points);
    * for a predicate, it is not the same for every point (a predicate that
      selects all or nothing is confirmed point by point).

    Otherwise ``fn`` is called point by point.
    """
    pts = np.asarray(pts, float)
    M = len(pts)
    extra = tuple(np.asarray(e) for e in extra)
    if M == 0:
        return np.zeros(0)
    try:
        out = np.asarray(fn(pts.T, *extra))
    except Exception:
        return _per_point(fn, pts, extra)
    if out.shape != (M,) or out.dtype == object:
        return _per_point(fn, pts, extra)
    rng = np.random.default_rng(0x5EED)
    pick = [0, M - 1]
    if out.dtype.kind == "b":
        if out.all() or not out.any():
            return _per_point(fn, pts, extra)
        for val in (True, False):
            idx = np.flatnonzero(out == val)
            pick.extend(rng.choice(idx, min(len(idx), _SAMPLE // 2), replace=False).tolist())
    else:
        pick.extend(rng.choice(M, min(M, _SAMPLE), replace=False).tolist())
    S = np.unique(np.asarray(pick, np.int64))
    far = 1e6 * (1.0 + np.abs(pts).max())
    ext = np.concatenate([pts, np.full((1, pts.shape[1]), far), np.full((1, pts.shape[1]), -far),
                          np.zeros((1, pts.shape[1]))])
    try:
        one = _per_point(fn, pts[S], tuple(e[S] for e in extra))
        sub = np.asarray(fn(pts[S].T, *(e[S] for e in extra)))
        aug = np.asarray(fn(ext.T, *(np.concatenate([e, e[:3]]) for e in extra)))
    except Exception:
        return _per_point(fn, pts, extra)
    if _same(out[S], one) and _same(out[S], sub) and aug.shape == (M + 3,) and _same(out, aug[:M]):
        return out
    return _per_point(fn, pts, extra)


def location_flags(fn, points, verb="got"):
    """Boolean (M,) flags of a one- or two-argument location function over
    the points (a two-argument one also receives the node index), evaluated
    per point (feax/fe.py:288-301, feax/DCboundary.py:448-460)."""
    nargs = fn.__code__.co_argcount if hasattr(fn, "__code__") else 1
    if nargs == 1:
        out = eval_points(fn, points)
    elif nargs == 2:
        out = eval_points(fn, points, np.arange(len(points)))
    else:
        raise ValueError(f"Wrong number of arguments for location_fn: must be 1 or 2, {verb} {nargs}")
    return np.asarray(out).reshape(len(points)).astype(bool)


def boundary_conditions_inds(points, cells, location_fns, ele_type="HEX8"):
    """List of (F, 2) int arrays (cell, local face) whose vertices all satisfy
    each location_fn, in np.argwhere (cell, face) order -- feax/fe.py:267-318.
    Two-argument location functions receive (point, node index)."""
    out = []
    if not location_fns:
        return out
    fverts = cells[:, element_info(ele_type)["faces"]]     # (E, faces, verts) node ids
    for fn in location_fns:
        flag = location_flags(fn, points, "get")
        on = flag[fverts].all(axis=2)
        out.append(np.argwhere(on).astype(np.int64))
    return out
