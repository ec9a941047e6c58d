"""The sparse Euclidean MST used for ridge sign propagation against the
reference's dense routine (scipy minimum_spanning_tree on pdist). CPU only."""
import numpy as np
from scipy.sparse.csgraph import minimum_spanning_tree
from scipy.spatial.distance import pdist, squareform

from miccai2016_amd.processing import ridge_graph
from oracle import fp64


def _tree_weight(pts, edges):
    return np.linalg.norm(pts[edges[:, 0]] - pts[edges[:, 1]], axis=1).sum()


def _is_spanning_tree(n, edges):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    g = coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n))
    return len(edges) == n - 1 and connected_components(g, directed=False)[0] == 1


def test_sparse_mst_equals_dense_without_ties(rng, monkeypatch):
    # continuous coordinates: distinct lengths, unique tree
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    pts = np.concatenate([rng.uniform(0, 12, size=(700, 3)),          # dense blob
                          rng.uniform(30, 60, size=(80, 3))])          # sparse far points
    edges = ridge_graph.spanning_tree_edges(pts, 1.5)
    dense = minimum_spanning_tree(squareform(pdist(pts))).tocoo()
    want = {tuple(sorted(e)) for e in zip(dense.row, dense.col)}
    got = {tuple(sorted(e)) for e in edges.tolist()}
    assert got == want
    dirs = rng.normal(size=(len(pts), 3)).astype(np.float32)
    np.testing.assert_array_equal(ridge_graph.propagate_signs(dirs, edges),
                                  fp64.mst_signs(pts, dirs, (1, 1, 1)))


def test_sparse_mst_on_lattice_has_minimum_weight(rng, monkeypatch):
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    grid = np.argwhere(rng.random((20, 20, 20)) < 0.15).astype(np.float64)
    edges = ridge_graph.spanning_tree_edges(grid, np.sqrt(3))
    assert _is_spanning_tree(len(grid), edges)
    dense = minimum_spanning_tree(squareform(pdist(grid))).tocoo()
    assert abs(_tree_weight(grid, edges) - dense.data.sum()) < 1e-6 * dense.data.sum()
    dirs = rng.normal(size=(len(grid), 3)).astype(np.float32)
    out = ridge_graph.propagate_signs(dirs, edges)
    np.testing.assert_array_equal(np.abs(out), np.abs(dirs))
    assert (np.einsum("ij,ij->i", out[edges[:, 0]], out[edges[:, 1]]) >= 0).all()


def test_large_ridge_is_fast(rng):
    import time
    grid = np.argwhere(rng.random((64, 64, 64)) < 0.1).astype(np.float64)   # ~26k points
    t = time.time()
    edges = ridge_graph.spanning_tree_edges(grid, np.sqrt(3))
    assert _is_spanning_tree(len(grid), edges)
    assert time.time() - t < 60
