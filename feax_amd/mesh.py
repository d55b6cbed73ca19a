"""Mesh container and the implicit structured HEX8 grid.

Mirrors feax.Mesh (feax/mesh.py:20-184, the container only) and
feax.StructuredGrid (feax/spgrid.py:28-186): same node / cell numbering and
the same HEX8 local corner order, so a structured box gives the same
(points, cells) arrays as in feax.
"""

import numpy as np

# feax HEX8 local corner offsets (dx, dy, dz) from a cell's base corner
# (feax/spgrid.py:21-25).
HEX8_OFFSETS = np.array([(0, 1, 1), (0, 0, 1), (0, 0, 0), (0, 1, 0),
                         (1, 1, 1), (1, 0, 1), (1, 0, 0), (1, 1, 0)], dtype=np.int64)


class Mesh:
    """points (N, dim) float64, cells (E, nodes_per_cell) int, ele_type."""

    def __init__(self, points, cells, ele_type="HEX8"):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells)
        self.ele_type = ele_type

    @property
    def num_nodes(self):
        return len(self.points)

    @property
    def num_cells(self):
        return len(self.cells)


class StructuredGrid:
    """Uniform HEX8 grid: cell (cx*ny + cy)*nz + cz, node (i*(ny+1) + j)*(nz+1) + k
    (feax/spgrid.py:28-41)."""

    ele_type = "HEX8"

    def __init__(self, dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
        self.nx, self.ny, self.nz = (int(d) for d in dims)
        self.spacing = np.asarray(spacing, float)
        self.origin = np.asarray(origin, float)

    @property
    def num_cells(self):
        return self.nx * self.ny * self.nz

    @property
    def num_nodes(self):
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    def cell_ijk(self, cell_ids):
        e = np.asarray(cell_ids, np.int64)
        nynz = self.ny * self.nz
        cx, r = e // nynz, e % nynz
        return cx, r // self.nz, r % self.nz

    def cell_to_nodes(self, cell_ids):
        """(n, 8) global node ids in feax HEX8 local order (spgrid.py:115-123)."""
        cx, cy, cz = self.cell_ijk(cell_ids)
        ny1, nz1 = self.ny + 1, self.nz + 1
        I = cx[:, None] + HEX8_OFFSETS[None, :, 0]
        J = cy[:, None] + HEX8_OFFSETS[None, :, 1]
        K = cz[:, None] + HEX8_OFFSETS[None, :, 2]
        return (I * ny1 + J) * nz1 + K

    def node_coords(self, node_ids):
        n = np.asarray(node_ids, np.int64)
        ny1, nz1 = self.ny + 1, self.nz + 1
        i, r = n // (ny1 * nz1), n % (ny1 * nz1)
        j, k = r // nz1, r % nz1
        return self.origin + np.stack([i, j, k], axis=-1) * self.spacing

    def node_id(self, i, j, k):
        i, j, k = (np.asarray(v, np.int64) for v in (i, j, k))
        return (i * (self.ny + 1) + j) * (self.nz + 1) + k

    def cell_id(self, ex, ey, ez):
        ex, ey, ez = (np.asarray(v, np.int64) for v in (ex, ey, ez))
        return (ex * self.ny + ey) * self.nz + ez

    def cell_centroids(self, cell_ids):
        cx, cy, cz = self.cell_ijk(cell_ids)
        return self.origin + (np.stack([cx, cy, cz], axis=-1) + 0.5) * self.spacing

    def to_mesh(self, cell_ids=None):
        """Explicit Mesh of ``cell_ids`` (default all), nodes renumbered by
        unique + searchsorted (feax/spgrid.py:156-170)."""
        ids = (np.arange(self.num_cells, dtype=np.int64) if cell_ids is None
               else np.asarray(cell_ids, np.int64).ravel())
        cells = self.cell_to_nodes(ids)
        used = np.unique(cells)
        return Mesh(self.node_coords(used), np.searchsorted(used, cells), ele_type=self.ele_type)


def rectangle_mesh(Nx, Ny, domain_x=1.0, domain_y=1.0, origin=(0.0, 0.0), ele_type="QUAD4"):
    """Structured QUAD4 mesh of a rectangle (feax/mesh.py:447-502): node
    i (Ny + 1) + j at (x_i, y_j), cell (i Ny + j) = counter-clockwise
    [(i, j), (i+1, j), (i+1, j+1), (i, j+1)]."""
    if ele_type != "QUAD4":
        raise NotImplementedError(f"rectangle_mesh builds QUAD4 here (got {ele_type})")
    x0, y0 = origin
    x = np.linspace(x0, x0 + domain_x, Nx + 1)
    y = np.linspace(y0, y0 + domain_y, Ny + 1)
    xv, yv = np.meshgrid(x, y, indexing="ij")
    points = np.stack([xv.reshape(-1), yv.reshape(-1)], axis=1)
    i, j = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    i, j = i.reshape(-1), j.reshape(-1)
    n0 = i * (Ny + 1) + j
    cells = np.stack([n0, n0 + (Ny + 1), n0 + (Ny + 1) + 1, n0 + 1], axis=1).astype(np.int32)
    return Mesh(points, cells, ele_type="QUAD4")


# HEX20 edge nodes (local 8..19): the midpoints of these corner pairs (the
# meshio / Abaqus hexahedron20 order, feax/basis.py:71-77).
HEX20_EDGES = np.array([(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                        (0, 4), (1, 5), (2, 6), (3, 7)], dtype=np.int64)


def hex20_mesh(dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """A structured HEX20 box (feax reads HEX20 meshes from files; this
    builds one): the corners of StructuredGrid(dims) with their node ids and
    HEX8 corner order, plus one node at the midpoint of every grid edge,
    numbered after the corners in the order the cells first list them."""
    grid = StructuredGrid(dims, spacing, origin)
    m8 = grid.to_mesh()
    N8 = len(m8.points)
    pairs = np.sort(m8.cells[:, HEX20_EDGES], axis=2).reshape(-1, 2)       # (12 E, 2)
    key = pairs[:, 0] * N8 + pairs[:, 1]
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    new_id = np.empty(len(uniq), np.int64)
    new_id[order] = np.arange(len(uniq))
    edge_nodes = (N8 + new_id[inv]).reshape(len(m8.cells), 12)
    firsts = pairs[first[order]]
    mids = 0.5 * (m8.points[firsts[:, 0]] + m8.points[firsts[:, 1]])
    return Mesh(np.concatenate([m8.points, mids]), np.concatenate([m8.cells, edge_nodes], axis=1),
                ele_type="HEX20")
