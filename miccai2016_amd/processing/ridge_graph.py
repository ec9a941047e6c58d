"""Minimum spanning tree over ridge voxels and sign propagation along it.

The reference builds the complete Euclidean graph of the ridge voxels
(pdist + squareform, adjust_directions.py:374-379) and runs scipy's
minimum_spanning_tree on it: O(R^2) time and memory, which is out of reach
for the tens of thousands of ridge voxels a noisy phantom produces. Here the
same tree is found from a sparse graph:

1. Kruskal (scipy's minimum_spanning_tree) on the edges no longer than
   `r0` (the scaled 26-neighbourhood). Every MST edge of length <= r0 is in
   this forest: such an edge is the lightest across a cut inside its
   component, and every edge leaving the component is longer than r0.
2. Boruvka rounds join the remaining components with their shortest
   outgoing edges, found by nearest-neighbour queries in a KD-tree.

With distinct edge lengths the result is the dense routine's tree; with
ties it is another minimum spanning tree of the same total length.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import breadth_first_order, connected_components, minimum_spanning_tree
from scipy.spatial import cKDTree
from scipy.spatial.distance import pdist, squareform

DENSE_LIMIT = 4000   # up to this many points the reference's dense routine is used


def _dense_mst(pts):
    return minimum_spanning_tree(squareform(pdist(pts, "euclidean")), overwrite=True).tocoo()


def _shortest_foreign_edges(pts, labels, ncomp, tree):
    """For every component, its shortest edge to another component:
    arrays (u, v, w) indexed by component (u = -1 if none)."""
    n = len(pts)
    best_w = np.full(ncomp, np.inf)
    best_u = np.full(ncomp, -1, np.int64)
    best_v = np.full(ncomp, -1, np.int64)
    # A component holding more than an eighth of the points (a long ridge) is
    # searched from its own points against a tree of all other points: its
    # points' nearest neighbours are mostly its own, so growing k would be
    # slow. The rest start with the shared tree and a growing k.
    sizes = np.bincount(labels, minlength=ncomp)
    big = sizes > max(n // 8, 64)
    pending = ~big[labels]
    for c in np.nonzero(big)[0]:
        _search_against_others(pts, labels, c, best_w, best_u, best_v)
    k = 8
    while pending.any() and k <= 64:
        idx = np.nonzero(pending)[0]
        kk = min(k, n)
        dist, nb = tree.query(pts[idx], k=kk)
        if kk == 1:
            dist, nb = dist[:, None], nb[:, None]
        foreign = labels[nb] != labels[idx][:, None]
        has = foreign.any(1)
        first = np.argmax(foreign, 1)
        rows = np.nonzero(has)[0]
        d = dist[rows, first[rows]]
        u = idx[rows]
        v = nb[rows, first[rows]]
        # per component the lightest candidate (ties: smallest u, then v)
        comp = labels[u]
        order = np.lexsort((v, u, d, comp))
        cs, at = np.unique(comp[order], return_index=True)
        r = order[at]
        better = d[r] < best_w[cs]
        cs, r = cs[better], r[better]
        best_w[cs], best_u[cs], best_v[cs] = d[r], u[r], v[r]
        pending[idx[has]] = False
        # a point whose k nearest are all in its own component may still hold
        # its component's shortest outgoing edge only if that edge is longer
        # than its k-th neighbour distance
        still = idx[~has]
        if len(still):
            kth = dist[~has, -1]
            needed = kth < best_w[labels[still]]
            pending[still[~needed]] = False
        k *= 2
    # points still pending: query against the points of other components only
    if pending.any():
        for c in np.unique(labels[pending]):
            _search_against_others(pts, labels, c, best_w, best_u, best_v,
                                   np.nonzero(pending & (labels == c))[0])
    return best_u, best_v, best_w


def _search_against_others(pts, labels, c, best_w, best_u, best_v, mine=None):
    """Shortest edge from component c (or its points `mine`) to any other
    component, by a KD-tree of the other components' points."""
    mine = np.nonzero(labels == c)[0] if mine is None else mine
    others = np.nonzero(labels != c)[0]
    if len(others) == 0 or len(mine) == 0:
        return
    d, j = cKDTree(pts[others]).query(pts[mine], k=1)
    r = int(np.argmin(d))
    if d[r] < best_w[c]:
        best_w[c], best_u[c], best_v[c] = d[r], mine[r], others[j[r]]


def spanning_tree_edges(pts, r0):
    """Edges (i, j) of a minimum spanning tree of the complete Euclidean graph
    of pts (R x 3 float64)."""
    n = len(pts)
    if n < 2:
        return np.zeros((0, 2), np.int64)
    if n <= DENSE_LIMIT:
        m = _dense_mst(pts)
        return np.stack([m.row, m.col], 1).astype(np.int64)
    tree = cKDTree(pts)
    pairs = tree.query_pairs(r0 * (1 + 1e-9), output_type="ndarray")
    if len(pairs):
        w = np.linalg.norm(pts[pairs[:, 0]] - pts[pairs[:, 1]], axis=1)
        g = coo_matrix((w, (pairs[:, 0], pairs[:, 1])), shape=(n, n)).tocsr()
        f = minimum_spanning_tree(g).tocoo()
        edges = [np.stack([f.row, f.col], 1).astype(np.int64)]
    else:
        edges = [np.zeros((0, 2), np.int64)]
    while True:
        e = np.concatenate(edges)
        g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
        ncomp, labels = connected_components(g, directed=False)
        if ncomp == 1:
            return e
        bu, bv, _ = _shortest_foreign_edges(pts, labels, ncomp, tree)
        ok = bu >= 0
        new = np.stack([bu[ok], bv[ok]], 1)
        # two components may pick the same edge; keep one copy
        new = np.unique(np.sort(new, 1), axis=0)
        # joining along several shortest edges at once can close a cycle when
        # lengths tie; drop edges that would join already-joined components
        keep = []
        parent = np.arange(ncomp)

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a
        w_new = np.linalg.norm(pts[new[:, 0]] - pts[new[:, 1]], axis=1)
        for idx in np.argsort(w_new, kind="stable"):
            a, b = find(labels[new[idx, 0]]), find(labels[new[idx, 1]])
            if a != b:
                parent[a] = b
                keep.append(idx)
        edges.append(new[np.asarray(keep, np.int64)])


def propagate_signs(dirs, edges, root=0):
    """Signs along the tree from `root`: a child keeps its direction if its
    dot product with its (already adjusted) parent is >= 0, else flips
    (adjust_directions.py:384-417). Returns the adjusted float32 directions.

    With s = dot(child, parent) of the unadjusted directions, the child's sign
    is the parent's times sign(s) when s != 0, +1 when s == 0 and -1 when s is
    NaN; the products along the root paths are formed by pointer jumping
    (log2(depth) vectorised rounds) instead of a walk over the nodes."""
    n = len(dirs)
    out = np.full((n, 3), np.nan, np.float32)
    if n == 0:
        return out
    d32 = np.asarray(dirs, np.float32)
    out[root] = d32[root]
    if len(edges) == 0:
        return out
    g = coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n)).tocsr()
    order, pred = breadth_first_order(g, root, directed=False, return_predecessors=True)
    nodes = order[1:]
    par = pred[nodes]
    dot = np.einsum("ij,ij->i", d32[nodes], d32[par], dtype=np.float32)
    val = np.ones(n, np.int8)
    nxt = np.arange(n)
    link = (dot != 0) & ~np.isnan(dot)       # sign = parent's sign x sign(dot)
    val[nodes[link]] = np.where(dot[link] > 0, 1, -1).astype(np.int8)
    nxt[nodes[link]] = par[link]
    val[nodes[np.isnan(dot)]] = -1           # the >= 0 test fails: flip
    while True:
        nn = nxt[nxt]
        if np.array_equal(nn, nxt):
            break
        val = val * val[nxt]
        nxt = nn
    val = val * val[nxt]                     # the last hop onto a fixed node
    out[order] = d32[order] * val[order].astype(np.float32)[:, None]
    return out
