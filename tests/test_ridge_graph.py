"""The sparse Euclidean MST used for ridge sign propagation against the
reference's dense routine (scipy minimum_spanning_tree on pdist). CPU only."""
import numpy as np
import pytest
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
    assert time.time() - t < 10


@pytest.mark.parametrize("density,seed", [(0.02, 1), (0.08, 2), (0.3, 3), (0.6, 4)])
def test_sparse_mst_weight_on_lattices(monkeypatch, density, seed):
    """Lattice points tie heavily; the sparse tree has the dense routine's
    total length, also with anisotropic voxel sizes and far-apart clusters."""
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    rng = np.random.default_rng(seed)
    grid = np.argwhere(rng.random((16, 16, 16)) < density).astype(np.float64)
    far = np.argwhere(rng.random((6, 6, 6)) < 0.3).astype(np.float64) + 40
    for vs in ((1, 1, 1), (2.0, 0.5, 1.25)):
        pts = np.concatenate([grid, far]) * np.asarray(vs)
        edges = ridge_graph.spanning_tree_edges(pts, float(np.linalg.norm(vs)))
        assert _is_spanning_tree(len(pts), edges)
        dense = minimum_spanning_tree(squareform(pdist(pts))).tocoo()
        assert abs(_tree_weight(pts, edges) - dense.data.sum()) <= 1e-9 * dense.data.sum()


def test_sparse_mst_is_the_dense_tree_for_distinct_lengths(rng, monkeypatch):
    """Continuous coordinates in clusters of very different sizes (one
    component holding most points): the unique tree, edge for edge."""
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    t = np.linspace(0, 6 * np.pi, 1500)
    helix = np.stack([t * 2, 10 * np.cos(t), 10 * np.sin(t)], 1) + rng.normal(0, 0.05, (1500, 3))
    pts = np.concatenate([helix, rng.uniform(-30, 60, size=(300, 3))])
    edges = ridge_graph.spanning_tree_edges(pts, 1.0)
    dense = minimum_spanning_tree(squareform(pdist(pts))).tocoo()
    assert {tuple(sorted(e)) for e in edges.tolist()} == \
        {tuple(sorted(e)) for e in zip(dense.row.tolist(), dense.col.tolist())}


def test_propagate_signs_zero_nan_and_forests():
    """dot >= 0 keeps (a zero product keeps), dot < 0 flips, a NaN product
    flips; the sign is taken against the adjusted parent
    (adjust_directions.py:384-417); a tree the root does not reach starts at
    its smallest index, which keeps its direction."""
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, -1, 0], [np.nan, 0, 0], [-1, 0, 0],
                     [-0.5, 0.5, 0], [3, 3, 3]], np.float32)
    edges = np.array([[0, 1], [2, 1], [3, 0], [4, 0], [1, 5]], np.int64)
    out = ridge_graph.propagate_signs(dirs, edges, 0)
    want = np.array([[1, 0, 0], [0, 1, 0], [0, 1, 0], [np.nan, -0.0, -0.0], [1, 0, 0],
                     [-0.5, 0.5, 0]], np.float32)
    np.testing.assert_array_equal(out[:6], want)
    assert np.signbit(out[3, 1]) and np.array_equal(out[6], dirs[6])
    # a second tree: 6 - 7 - 8, started at 6
    dirs2 = np.concatenate([dirs, [[-3, -3, 0], [1, 1, 0]]]).astype(np.float32)
    edges2 = np.concatenate([edges, [[7, 6], [7, 8]]])
    out2 = ridge_graph.propagate_signs(dirs2, edges2, 0)
    np.testing.assert_array_equal(out2[6:], [[3, 3, 3], [3, 3, 0], [1, 1, 0]])


class _NoDevice:
    """Stand-in loader for the adjuster's host steps."""


def _segments(rng, shape, count, pair=None):
    """Random 26-connected ridge segments (random walks) and random
    directions; with pair "zero", one segment is two neighbours with
    orthogonal axis directions, with "tiny" two neighbours whose float32 dot
    product is 1e-7 (small, but its sign flips with the parent's); with
    "straddle", one segment is a straight chain of 40 voxels whose halves
    point along two orthogonal axes (one zero product, pieces of 20)."""
    ridge = np.zeros(shape, np.float32)
    for _ in range(count):
        p = rng.integers(2, np.array(shape) - 2)
        for _ in range(int(rng.integers(5, 60))):
            ridge[tuple(p)] = 1
            p = np.clip(p + rng.integers(-1, 2, size=3), 0, np.array(shape) - 1)
    dirs = rng.normal(size=shape + (3,)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    if pair in ("zero", "tiny"):      # a segment of two voxels, one forest edge
        ridge[:4, :4, :4] = 0
        ridge[1, 1, 1] = ridge[1, 1, 2] = 1
        dirs[1, 1, 1] = (1, 0, 0)
        dirs[1, 1, 2] = (0, 1, 0) if pair == "zero" else (1e-7, 1, 0)
    if pair == "straddle":
        ridge[:, :4, :4] = 0
        ridge[5:45, 1, 1] = 1
        dirs[5:25, 1, 1] = (0, 0, 1)
        dirs[25:45, 1, 1] = (0, 1, 0)
    return ridge, dirs


@pytest.mark.parametrize("seed,pair", [(1, None), (2, None), (3, "zero"), (4, "tiny"),
                                       (5, "straddle")])
def test_cleaning_with_forest_signs_equals_full_tree(monkeypatch, seed, pair):
    """The cleaning keeps the same segments whether its scores come from the
    neighbour forest's signs or from the full spanning tree's. A tiny forest
    product keeps the forest's signs; an exactly zero one leaves its pieces'
    relative sign open, which the score bounds settle for a short segment
    (no tree) and not for one whose bounds straddle the cut (the tree)."""
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    rng = np.random.default_rng(seed)
    shape = (64, 64, 64)
    ridge, dirs = _segments(rng, shape, 80, pair)
    adj = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size=1, loader=_NoDevice())
    c_full, d_full = adj._adjust_ridge_directions(ridge, dirs)
    trees = []
    tree_edges = ridge_graph.ridge_tree_edges
    monkeypatch.setattr(ridge_graph, "ridge_tree_edges",
                        lambda *a: trees.append(1) or tree_edges(*a))
    c_forest, d_forest, parts = adj._adjust_ridge_directions(ridge, dirs, for_cleaning=True)
    assert (parts is None) == (pair not in ("zero", "straddle"))
    np.testing.assert_array_equal(c_full, c_forest)
    np.testing.assert_array_equal(np.abs(d_full), np.abs(d_forest))
    keep_full = adj._do_clean_ridge(ridge, c_full, d_full)
    keep_forest = adj._do_clean_ridge(ridge, c_forest, d_forest, parts)
    assert len(trees) == (1 if pair == "straddle" else 0)
    assert 0 < keep_full.sum() < ridge.sum()
    np.testing.assert_array_equal(keep_full, keep_forest)


def _forest_components(n, forest):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    f = np.asarray(forest).reshape(-1, 2)
    g = coo_matrix((np.ones(len(f)), (f[:, 0], f[:, 1])), shape=(n, n))
    return connected_components(g, directed=False)


@pytest.mark.parametrize("vs", [(1, 1, 1), (2.0, 0.5, 1.25), (1, 1, 3)])
@pytest.mark.parametrize("seed", [0, 1])
def test_tree_from_forest_equals_ridge_tree(vs, seed):
    """mh_ridge_tree_from_forest_host, given the neighbour forest of the
    points or of a superset restricted to whole trees of it, returns
    mh_ridge_tree_host's edges, in order."""
    rng = np.random.default_rng(seed)
    shape = (40, 40, 40)
    ridge, _ = _segments(rng, shape, 60)
    coords = np.argwhere(ridge != 0)
    forest = ridge_graph.ridge_forest_edges(coords, shape, vs)
    np.testing.assert_array_equal(ridge_graph.ridge_tree_from_forest(coords, shape, vs, forest),
                                  ridge_graph.ridge_tree_edges(coords, shape, vs))
    ncomp, comp = _forest_components(len(coords), forest)
    assert ncomp > 4
    keep = (rng.random(ncomp) < 0.6)[comp]
    sub = ridge_graph.restrict_forest(forest, keep)
    assert len(coords[keep]) > ridge_graph.DENSE_LIMIT
    np.testing.assert_array_equal(
        ridge_graph.ridge_tree_from_forest(coords[keep], shape, vs, sub),
        ridge_graph.ridge_tree_edges(coords[keep], shape, vs))


def test_tree_from_forest_rejects_other_forests(monkeypatch):
    """A forest that is not a forest of short neighbour edges of the points
    is refused; restrict_forest refuses a mask that splits a tree."""
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 1)
    shape = (8, 8, 8)
    coords = np.array([[1, 1, 1], [1, 1, 2], [1, 2, 2], [5, 5, 5]])
    vs = (1, 1, 1)
    forest = ridge_graph.ridge_forest_edges(coords, shape, vs)
    assert forest.tolist() == [[0, 1], [1, 2]]
    want = ridge_graph.ridge_tree_edges(coords, shape, vs)
    np.testing.assert_array_equal(ridge_graph.ridge_tree_from_forest(coords, shape, vs, forest), want)
    np.testing.assert_array_equal(
        ridge_graph.ridge_tree_from_forest(coords, shape, vs, np.zeros((0, 2), np.int64)), want)
    for bad in ([[0, 1], [1, 2], [0, 2]],     # more edges than a forest holds
                [[0, 1], [0, 1]],             # a cycle
                [[0, 3]],                     # a long edge
                [[0, 4]], [[-1, 0]]):         # out of range
        with pytest.raises(RuntimeError):
            ridge_graph.ridge_tree_from_forest(coords, shape, vs, np.array(bad))
    with pytest.raises(RuntimeError):         # a voxel outside the volume
        ridge_graph.ridge_tree_from_forest(coords, (4, 8, 8), vs, forest)
    with pytest.raises(ValueError):
        ridge_graph.restrict_forest(forest, [True, False, True, True])
    np.testing.assert_array_equal(ridge_graph.restrict_forest(forest, [True, True, True, False]),
                                  [[0, 1], [1, 2]])
    np.testing.assert_array_equal(
        ridge_graph.restrict_forest([[1, 3], [3, 4]], [False, True, False, True, True]),
        [[0, 1], [1, 2]])


@pytest.mark.parametrize("seed", [1, 2, 6])
def test_final_tree_from_the_cleaning_forest(monkeypatch, seed):
    """The numpy API's adjustment builds the cleaned ridge's tree from the
    cleaning's neighbour forest, restricted to the kept segments: the same
    coordinates and signed directions, bit for bit, as building it anew."""
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    rng = np.random.default_rng(seed)
    shape = (64, 64, 64)
    ridge, dirs = _segments(rng, shape, 80)
    adj = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size=1, loader=_NoDevice())
    c, d, parts = adj._adjust_ridge_directions(ridge, dirs, for_cleaning=True)
    assert adj._forest is not None
    cleaned = adj._do_clean_ridge(ridge, c, d, parts)
    assert 0 < cleaned.sum() < ridge.sum()
    assert len(adj._forest[0]) == cleaned.sum()
    trees = []
    tree_edges = ridge_graph.ridge_tree_edges
    monkeypatch.setattr(ridge_graph, "ridge_tree_edges",
                        lambda *a: trees.append(1) or tree_edges(*a))
    c2, d2 = adj._adjust_ridge_directions(cleaned, dirs)
    assert trees == [] and adj._forest is None
    fresh = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size=1, loader=_NoDevice())
    c3, d3 = fresh._adjust_ridge_directions(cleaned, dirs)
    assert trees == [1]
    np.testing.assert_array_equal(c2, c3)
    np.testing.assert_array_equal(d2.view(np.uint32), d3.view(np.uint32))


@pytest.mark.parametrize("vs", [(1, 1, 1), (2.0, 0.5, 1.25), (1, 1, 3)])
def test_lattice_tree_is_the_point_tree(monkeypatch, vs):
    """The dense-lookup lattice construction gives the k-d tree
    construction's tree on the scaled points, edge for edge (both break ties
    on the index pair), with the dense routine's total length."""
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 10)
    rng = np.random.default_rng(7)
    shape = (20, 18, 22)
    mask = rng.random(shape) < 0.12
    mask[2:18, 9, 11] = True
    coords = np.argwhere(mask)
    pts = coords * np.asarray(vs, np.float64)
    a = ridge_graph.ridge_tree_edges(coords, shape, vs)
    b = ridge_graph.spanning_tree_edges(pts, float(np.linalg.norm(vs)))
    assert {tuple(sorted(e)) for e in a.tolist()} == {tuple(sorted(e)) for e in b.tolist()}
    dense = minimum_spanning_tree(squareform(pdist(pts))).tocoo()
    assert abs(_tree_weight(pts, a) - dense.data.sum()) <= 1e-9 * dense.data.sum()


def test_lattice_tree_rejects_bad_coordinates(monkeypatch):
    monkeypatch.setattr(ridge_graph, "DENSE_LIMIT", 1)
    with pytest.raises(RuntimeError):
        ridge_graph.ridge_tree_edges(np.array([[0, 0, 0], [0, 0, 0]]), (2, 2, 2), 1)
    with pytest.raises(RuntimeError):
        ridge_graph.ridge_tree_edges(np.array([[0, 0, 0], [0, 0, 2]]), (2, 2, 2), 1)


@pytest.mark.parametrize("seed", range(12))
def test_kept_segments_bounds_hold_for_every_piece_sign(seed):
    """When the piece bounds decide the cleaning, the kept segments are the
    top ceil(3 ln(e + n)) (stable order) under every choice of the pieces'
    relative signs, checked by enumerating them; when some choice would
    change the set, the bounds must not decide."""
    import itertools
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    rng = np.random.default_rng(seed)
    adj = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size=1, loader=_NoDevice())
    ncomp = 14
    npieces = rng.integers(1, 4, size=ncomp)
    npieces[rng.choice(ncomp, size=ncomp - 4, replace=False)] = 1   # at most 4 open segments
    seg_of, vecs = [], []
    for s_, k in enumerate(npieces):
        scale = 10 ** rng.uniform(-1, 2)
        for j in range(k):            # open segments: a leading piece, smaller others
            seg_of.append(s_)
            vecs.append(rng.normal(size=3) * scale * (1.0 if j == 0 else rng.uniform(0.01, 0.6)))
    lab_of, vecs = np.array(seg_of), np.array(vecs)
    parts = np.arange(len(vecs))                 # one point per piece
    keep = int(np.ceil(3 * np.log(np.e + ncomp)))
    open_pieces = [i for i in range(len(vecs)) if npieces[lab_of[i]] > 1
                   and i != np.nonzero(lab_of == lab_of[i])[0][0]]
    sets = set()
    for flips in itertools.product((1.0, -1.0), repeat=len(open_pieces)):
        v = vecs.copy()
        v[open_pieces] *= np.array(flips)[:, None]
        lengths = np.linalg.norm(np.stack([np.bincount(lab_of, v[:, k], ncomp)
                                           for k in range(3)], 1), axis=1)
        order = np.argsort(-lengths, kind="stable")
        sets.add(tuple(sorted(order[:keep])))
    got = adj._kept_segments(lab_of, ncomp, vecs, parts)
    if got is not None:
        assert sets == {tuple(np.nonzero(got)[0])}
    # the bounds decide whenever the set cannot change, up to their slack:
    # with no open segment they always do
    if len(open_pieces) == 0:
        assert got is not None


@pytest.mark.parametrize("seed,zeros,nans", [(1, 0, 0), (2, 5, 0), (3, 40, 0), (4, 3, 1)])
def test_forest_pieces_against_the_walk_and_scipy(seed, zeros, nans):
    """The union-find pieces and signs: without zero products the signs are
    the breadth-first walk's, bit for bit; every non-zero edge
    joins points with non-negative signed products; the pieces are scipy's
    connected components of the forest without its zero-product edges, in
    order of their smallest point; zero and NaN products are counted."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    rng = np.random.default_rng(seed)
    n = 3000
    # a random forest: each point links to an earlier one, except tree roots
    parent = np.array([-1] + [int(rng.integers(0, i)) if rng.random() > 0.02 else -1
                              for i in range(1, n)])
    child = np.nonzero(parent >= 0)[0]
    edges = np.stack([parent[child], child], 1)
    rng.shuffle(edges)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    for k in rng.choice(len(edges), size=zeros + nans, replace=False)[:zeros]:
        a, b = edges[k]
        dirs[b] = np.cross(dirs[a], rng.normal(size=3)).astype(np.float32)
        dirs[b] -= dirs[a] * (np.dot(dirs[a], dirs[b]) / np.dot(dirs[a], dirs[a]))
        dirs[b] = (0.0, 0.0, 0.0) if rng.random() < 0.5 else dirs[b]
    if nans:
        dirs[edges[rng.integers(len(edges)), 1]] = (np.nan, 0.0, 0.0)
    a32, b32 = dirs[edges[:, 0]], dirs[edges[:, 1]]
    dots = (a32[:, 0] * b32[:, 0] + a32[:, 1] * b32[:, 1]) + a32[:, 2] * b32[:, 2]
    signs, pieces, nzero, nnan = ridge_graph.forest_pieces(dirs, edges)
    assert nzero == int((dots == 0).sum()) and nnan == int(np.isnan(dots).sum())
    keep = ~((dots == 0) | np.isnan(dots))
    e = edges[keep]
    ncc, lab = connected_components(coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])),
                                               shape=(n, n)), directed=False)
    # scipy's labels renumbered by smallest point equal ours
    first = np.full(ncc, n)
    np.minimum.at(first, lab, np.arange(n))
    order = np.argsort(first)
    renum = np.empty(ncc, np.int64)
    renum[order] = np.arange(ncc)
    np.testing.assert_array_equal(pieces, renum[lab])
    s = signs[e[:, 0]], signs[e[:, 1]]
    prod = (s[0][:, 0] * s[1][:, 0] + s[0][:, 1] * s[1][:, 1]) + s[0][:, 2] * s[1][:, 2]
    assert (prod >= 0).all()
    np.testing.assert_array_equal(np.abs(signs), np.abs(dirs))
    if zeros == 0 and nans == 0:
        # every tree's smallest point keeps its direction in both (the walk
        # starts the tree of point 0 there), so the signs are the walk's bit
        # for bit
        walk = ridge_graph.propagate_signs(dirs, edges, 0)
        np.testing.assert_array_equal(walk.view(np.uint32), signs.view(np.uint32))


def _walk_signs(dirs, edges, root):
    """The reference's traversal restated in plain Python: breadth first
    from root, then from every unreached point in index order; a child keeps
    its direction when its float32 product ((c0 p0 + c1 p1) + c2 p2) with
    its adjusted parent is >= 0, else flips (a NaN product flips)."""
    from collections import deque
    n = len(dirs)
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    out = np.array(dirs, np.float32)
    seen = np.zeros(n, bool)

    def grow(r):
        seen[r] = True
        q = deque([r])
        while q:
            p = q.popleft()
            for c in adj[p]:
                if seen[c]:
                    continue
                seen[c] = True
                cd, pp = dirs[c], out[p]
                d = (cd[0] * pp[0] + cd[1] * pp[1]) + cd[2] * pp[2]
                out[c] = cd * np.float32(1.0 if d >= 0 else -1.0)   # a sign times the vector
                q.append(c)
    grow(root)
    for i in range(n):
        if not seen[i]:
            grow(i)
    return out


@pytest.mark.parametrize("seed,cuts,root", [(1, 0, 0), (2, 12, 0), (3, 60, 57), (4, 20, 1999)])
def test_tree_signs_equal_the_walk(seed, cuts, root):
    """mh_tree_signs_host (pieces between zero and NaN products, anchored by
    the walk over the pieces) gives the reference traversal's directions bit
    for bit on forests with zero and NaN products and several trees."""
    rng = np.random.default_rng(seed)
    n = 2000
    parent = np.array([-1] + [int(rng.integers(0, i)) if rng.random() > 0.01 else -1
                              for i in range(1, n)])
    child = np.nonzero(parent >= 0)[0]
    edges = np.stack([parent[child], child], 1)
    rng.shuffle(edges)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    for k, idx in enumerate(rng.choice(len(edges), size=cuts, replace=False)):
        b = edges[idx, 1]
        if k % 3 == 0:
            dirs[b] = (0.0, 0.0, 0.0)                     # zero products
        elif k % 3 == 1:
            a = edges[idx, 0]
            v = np.zeros(3, np.float32)
            j = int(np.argmax(np.abs(dirs[a])))
            v[(j + 1) % 3] = 1.0                          # exactly orthogonal to an axis
            dirs[a] = 0.0
            dirs[a, j] = 1.0
            dirs[b] = v
        else:
            dirs[b, 0] = np.nan                           # NaN products
    got = ridge_graph.propagate_signs(dirs, edges, root)
    want = _walk_signs(dirs, edges, root)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
