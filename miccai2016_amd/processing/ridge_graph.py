"""Minimum spanning tree over ridge voxels and sign propagation along it.

The reference builds the complete Euclidean graph of the ridge voxels
(pdist + squareform, adjust_directions.py:374-379) and runs scipy's
minimum_spanning_tree on it: O(R^2) time and memory, which is out of reach
for the tens of thousands of ridge voxels a noisy phantom produces. Small
ridges (up to DENSE_LIMIT points) still take that routine, so their ties in
length resolve as the reference's do. Larger ones take the library's host
C++ construction (mh_spanning_tree_host): Kruskal on the edges no longer than
the neighbourhood radius, then Boruvka rounds by k-d tree search. With
distinct edge lengths both give the same tree; with ties the sparse one is
another minimum spanning tree of the same total length.
"""
import numpy as np
from scipy.sparse.csgraph import minimum_spanning_tree
from scipy.spatial.distance import pdist, squareform

from .. import _lib

DENSE_LIMIT = 256    # up to this many points the reference's dense routine is used


def _dense_mst(pts):
    return minimum_spanning_tree(squareform(pdist(pts, "euclidean")), overwrite=True).tocoo()


def spanning_tree_edges(pts, r0):
    """Edges (i, j) of a minimum spanning tree of the complete Euclidean graph
    of pts (R x 3 float64); r0: the neighbourhood radius (scaled voxel
    diagonal)."""
    n = len(pts)
    if n < 2:
        return np.zeros((0, 2), np.int64)
    if n <= DENSE_LIMIT:
        m = _dense_mst(pts)
        return np.stack([m.row, m.col], 1).astype(np.int64)
    p = np.require(pts, np.float64, ["C"])
    edges = np.empty((n - 1, 2), np.int64)
    rc = _lib.lib().mh_spanning_tree_host(p.ctypes.data, n, float(r0), edges.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mh_spanning_tree_host failed (status {rc}): non-finite coordinates?")
    return edges


def ridge_tree_edges(coords, shape, voxel_size):
    """spanning_tree_edges of the voxels `coords` (R x 3 indices into
    `shape`) at positions coords * voxel_size, with the scaled voxel diagonal
    as the neighbourhood radius."""
    coords = np.asarray(coords)
    n = len(coords)
    if n <= DENSE_LIMIT:
        pts = np.multiply(coords, voxel_size).astype(np.float64)
        return spanning_tree_edges(pts, float(np.linalg.norm(voxel_size)))
    c = np.require(coords, np.int64, ["C"])
    vs = np.require(np.multiply(np.ones(3), voxel_size), np.float64, ["C"])
    edges = np.empty((n - 1, 2), np.int64)
    rc = _lib.lib().mh_ridge_tree_host(c.ctypes.data, n, _lib.dims3(shape), vs.ctypes.data,
                                       edges.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mh_ridge_tree_host failed (status {rc})")
    return edges


def ridge_forest_edges(coords, shape, voxel_size):
    """The first phase of ridge_tree_edges alone: the spanning forest of the
    neighbour pairs shorter than twice the smallest voxel size (for an
    isotropic voxel size, a spanning tree of every 26-connected segment)."""
    coords = np.asarray(coords)
    n = len(coords)
    if n < 2:
        return np.zeros((0, 2), np.int64)
    c = np.require(coords, np.int64, ["C"])
    vs = np.require(np.multiply(np.ones(3), voxel_size), np.float64, ["C"])
    edges = np.empty((n - 1, 2), np.int64)
    ne = _lib.lib().mh_ridge_forest_host(c.ctypes.data, n, _lib.dims3(shape), vs.ctypes.data,
                                         edges.ctypes.data)
    if ne < 0:
        raise RuntimeError(f"mh_ridge_forest_host failed (status {ne})")
    return edges[:ne]


def ridge_tree_from_forest(coords, shape, voxel_size, forest):
    """ridge_tree_edges(coords, shape, voxel_size) from its first phase:
    `forest` is ridge_forest_edges of these points, or of a superset
    restricted to a union of its trees (whole 26-connected segments, for an
    isotropic voxel size) and renumbered in order (restrict_forest). The
    edges equal ridge_tree_edges' (mh_ridge_tree_from_forest_host)."""
    coords = np.asarray(coords)
    n = len(coords)
    if n <= DENSE_LIMIT:
        return ridge_tree_edges(coords, shape, voxel_size)
    c = np.require(coords, np.int64, ["C"])
    vs = np.require(np.multiply(np.ones(3), voxel_size), np.float64, ["C"])
    f = np.require(forest, np.int64, ["C"]).reshape(-1, 2)
    edges = np.empty((n - 1, 2), np.int64)
    rc = _lib.lib().mh_ridge_tree_from_forest_host(c.ctypes.data, n, _lib.dims3(shape),
                                                   vs.ctypes.data, f.ctypes.data, len(f),
                                                   edges.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mh_ridge_tree_from_forest_host failed (status {rc})")
    return edges


def restrict_forest(forest, keep):
    """The edges of `forest` (E x 2) between points kept by the boolean mask
    `keep`, renumbered to the kept points' positions in order. Kept points
    must hold whole trees of the forest (an edge with one end kept raises)."""
    keep = np.asarray(keep, bool)
    f = np.asarray(forest, np.int64).reshape(-1, 2)
    ka, kb = keep[f[:, 0]], keep[f[:, 1]]
    if np.any(ka != kb):
        raise ValueError("restrict_forest: the kept points split a tree of the forest")
    new = np.cumsum(keep, dtype=np.int64) - 1
    return new[f[ka]]


def forest_pieces(dirs, edges):
    """Pieces of a forest between its zero-product edges and signs within
    them (mh_forest_pieces_host): (signed float32 directions, int32 piece
    labels 0 .. P-1, number of zero products, number of NaN products). In a
    forest without zero or NaN products the signs are propagate_signs(dirs,
    edges, 0)'s bit for bit: each tree's smallest point keeps its direction
    and every other sign is the parity of its path."""
    d32 = np.require(dirs, np.float32, ["C"]).reshape(-1, 3)
    n = len(d32)
    out = np.empty((n, 3), np.float32)
    pieces = np.empty(n, np.int32)
    counts = np.zeros(2, np.int64)
    e = np.require(edges, np.int64, ["C"]).reshape(-1, 2)
    rc = _lib.lib().mh_forest_pieces_host(d32.ctypes.data, n, e.ctypes.data, len(e),
                                          out.ctypes.data, pieces.ctypes.data,
                                          counts.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mh_forest_pieces_host failed (status {rc})")
    return out, pieces, int(counts[0]), int(counts[1])


def propagate_signs(dirs, edges, root=0):
    """Signs along the tree from `root`: a child keeps its direction if its
    dot product with its (already adjusted) parent is >= 0, else flips
    (adjust_directions.py:384-417). Returns the adjusted float32 directions.
    For a forest, every tree not reached from `root` starts at its smallest
    index, which keeps its direction."""
    d32 = np.require(dirs, np.float32, ["C"]).reshape(-1, 3)
    n = len(d32)
    out = np.empty((n, 3), np.float32)
    if n == 0:
        return out
    e = np.require(edges, np.int64, ["C"]).reshape(-1, 2)
    rc = _lib.lib().mh_tree_signs_host(d32.ctypes.data, n, e.ctypes.data, len(e), int(root),
                                       out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mh_tree_signs_host failed (status {rc})")
    return out
