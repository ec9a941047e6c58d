// Host C++: the minimum spanning tree of the complete Euclidean graph over
// ridge voxels, and sign propagation along it. Replaces the dense pdist +
// minimum_spanning_tree + depth-first walk of _adjust_ridge_directions
// (processing/adjust_directions.py:365-417), which costs O(R^2) time and
// memory, by an exact sparse construction:
//
// 1. Kruskal on the edges no longer than r0 (the scaled 26-neighbourhood).
//    Every tree edge of length <= r0 is in this forest: such an edge is the
//    lightest across a cut inside its component, and every edge leaving the
//    component is longer than r0.
// 2. Boruvka rounds join the remaining components, each along its shortest
//    edge to another component, found by a k-d tree search that skips
//    subtrees whose points all lie in the searching component.
//
// Edges compare by (length^2, smaller index, larger index), so the tree is
// one fixed minimum spanning tree whatever the visiting order; with distinct
// lengths it is the unique one. No GPU code here.
#include "mh_common.h"

#include <atomic>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

namespace {

struct Edge {
    double w2;
    int64_t a, b;  // a < b
};

inline bool lighter(const Edge &x, const Edge &y)
{
    if (x.w2 != y.w2) return x.w2 < y.w2;
    if (x.a != y.a) return x.a < y.a;
    return x.b < y.b;
}

struct DisjointSets {
    std::vector<int64_t> parent;
    explicit DisjointSets(int64_t n) : parent(size_t(n)) { std::iota(parent.begin(), parent.end(), 0); }
    int64_t find(int64_t x)
    {
        while (parent[size_t(x)] != x) {
            parent[size_t(x)] = parent[size_t(parent[size_t(x)])];
            x = parent[size_t(x)];
        }
        return x;
    }
    bool unite(int64_t x, int64_t y)
    {
        x = find(x);
        y = find(y);
        if (x == y) return false;
        if (x < y) std::swap(x, y);
        parent[size_t(x)] = y;
        return true;
    }
};

// k-d tree over the points: nodes hold a contiguous range of `order`, their
// bounding box and, during a Boruvka round, the component shared by all their
// points (-1 when mixed).
class KdTree {
public:
    static constexpr int kLeaf = 8;
    struct Node {
        double lo[3], hi[3];
        int64_t begin, end;
        int64_t left, right;  // -1 for leaves
        int64_t comp;
    };

    KdTree(const double *pts, int64_t n) : p_(pts), order_(size_t(n))
    {
        std::iota(order_.begin(), order_.end(), 0);
        nodes_.reserve(size_t(2 * (n / kLeaf + 1)));
        build(0, n);
    }

    // every pair (i, j), i < j, with |pi - pj|^2 <= r2
    void pairs_within(double r2, std::vector<Edge> &out) const
    {
        for (int64_t i = 0; i < int64_t(order_.size()); ++i) pairs_of(0, i, r2, out);
    }

    // component of every node from the point components
    void label(const std::vector<int64_t> &comp) { label_node(0, comp); }

    // shortest edge from point i (component c) to a point of another
    // component, if lighter than `best`
    void nearest_foreign(int64_t i, int64_t c, const std::vector<int64_t> &comp, Edge &best) const
    {
        foreign(0, i, c, comp, best);
    }

    int64_t size() const { return int64_t(order_.size()); }

private:
    const double *p_;
    std::vector<int64_t> order_;
    std::vector<Node> nodes_;

    const double *pt(int64_t i) const { return p_ + 3 * i; }

    int64_t build(int64_t b, int64_t e)
    {
        const int64_t id = int64_t(nodes_.size());
        nodes_.push_back(Node{});
        Node nd{};
        nd.begin = b;
        nd.end = e;
        nd.left = nd.right = -1;
        nd.comp = -1;
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = HUGE_VAL;
            nd.hi[k] = -HUGE_VAL;
        }
        for (int64_t t = b; t < e; ++t)
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::min(nd.lo[k], pt(order_[size_t(t)])[k]);
                nd.hi[k] = std::max(nd.hi[k], pt(order_[size_t(t)])[k]);
            }
        if (e - b > kLeaf) {
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (nd.hi[k] - nd.lo[k] > nd.hi[axis] - nd.lo[axis]) axis = k;
            const int64_t m = b + (e - b) / 2;
            std::nth_element(order_.begin() + b, order_.begin() + m, order_.begin() + e,
                             [&](int64_t x, int64_t y) {
                                 const double px = pt(x)[axis], py = pt(y)[axis];
                                 return px < py || (px == py && x < y);
                             });
            nd.left = build(b, m);
            nd.right = build(m, e);
        }
        nodes_[size_t(id)] = nd;
        return id;
    }

    double box_dist2(const Node &nd, const double *q) const
    {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = q[k] < nd.lo[k] ? nd.lo[k] - q[k] : (q[k] > nd.hi[k] ? q[k] - nd.hi[k] : 0.0);
            s += d * d;
        }
        return s;
    }

    double dist2(int64_t i, int64_t j) const
    {
        const double *a = pt(i), *b = pt(j);
        const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        return d0 * d0 + d1 * d1 + d2 * d2;
    }

    void pairs_of(int64_t id, int64_t i, double r2, std::vector<Edge> &out) const
    {
        const Node &nd = nodes_[size_t(id)];
        if (box_dist2(nd, pt(i)) > r2) return;
        if (nd.left < 0) {
            for (int64_t t = nd.begin; t < nd.end; ++t) {
                const int64_t j = order_[size_t(t)];
                if (j > i) {
                    const double d = dist2(i, j);
                    if (d <= r2) out.push_back(Edge{d, i, j});
                }
            }
            return;
        }
        pairs_of(nd.left, i, r2, out);
        pairs_of(nd.right, i, r2, out);
    }

    int64_t label_node(int64_t id, const std::vector<int64_t> &comp)
    {
        Node &nd = nodes_[size_t(id)];
        if (nd.left < 0) {
            int64_t c = comp[size_t(order_[size_t(nd.begin)])];
            for (int64_t t = nd.begin + 1; t < nd.end && c >= 0; ++t)
                if (comp[size_t(order_[size_t(t)])] != c) c = -1;
            nd.comp = c;
        } else {
            const int64_t a = label_node(nd.left, comp);
            const int64_t b = label_node(nd.right, comp);
            nodes_[size_t(id)].comp = (a == b) ? a : -1;
        }
        return nodes_[size_t(id)].comp;
    }

    void foreign(int64_t id, int64_t i, int64_t c, const std::vector<int64_t> &comp, Edge &best) const
    {
        const Node &nd = nodes_[size_t(id)];
        if (nd.comp == c) return;
        // strict: a box at exactly the best distance may hold a tie that
        // wins on the indices
        if (box_dist2(nd, pt(i)) > best.w2) return;
        if (nd.left < 0) {
            for (int64_t t = nd.begin; t < nd.end; ++t) {
                const int64_t j = order_[size_t(t)];
                if (comp[size_t(j)] == c) continue;
                const Edge e{dist2(i, j), std::min(i, j), std::max(i, j)};
                if (lighter(e, best)) best = e;
            }
            return;
        }
        const double dl = box_dist2(nodes_[size_t(nd.left)], pt(i));
        const double dr = box_dist2(nodes_[size_t(nd.right)], pt(i));
        if (dl <= dr) {
            foreign(nd.left, i, c, comp, best);
            foreign(nd.right, i, c, comp, best);
        } else {
            foreign(nd.right, i, c, comp, best);
            foreign(nd.left, i, c, comp, best);
        }
    }
};

// Kruskal over `shorts` (sorted here), extending the forest in `sets`.
void kruskal(std::vector<Edge> &shorts, DisjointSets &sets, int64_t *edges, int64_t &ne)
{
    std::sort(shorts.begin(), shorts.end(), lighter);
    for (const Edge &e : shorts)
        if (sets.unite(e.a, e.b)) {
            edges[2 * ne] = e.a;
            edges[2 * ne + 1] = e.b;
            ++ne;
        }
}

// Boruvka rounds that join the forest's components into one tree.
//
// Per point, a lower bound on its distance^2 to the nearest point of
// another component. Components only merge, so the set of foreign points
// only shrinks and a bound stays valid in later rounds. A component's
// points are searched in increasing bound order, and the search stops
// once the bound exceeds the component's best edge so far.
int complete_tree(KdTree &tree, int64_t n, DisjointSets &sets, int64_t *edges, int64_t ne)
{
    std::vector<int64_t> comp(static_cast<size_t>(n)), root_ix(static_cast<size_t>(n), -1);
    std::vector<int64_t> members(static_cast<size_t>(n)), first;
    std::vector<double> lb(static_cast<size_t>(n), 0.0);
    std::vector<Edge> best, picks;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    while (ne < n - 1) {
        int64_t ncomp = 0;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t r = sets.find(i);
            if (root_ix[size_t(r)] < 0) root_ix[size_t(r)] = ncomp++;
            comp[size_t(i)] = root_ix[size_t(r)];
        }
        for (int64_t i = 0; i < n; ++i) root_ix[size_t(sets.find(i))] = -1;
        tree.label(comp);
        // members of each component, by increasing bound (then index)
        first.assign(size_t(ncomp) + 1, 0);
        for (int64_t i = 0; i < n; ++i) ++first[size_t(comp[size_t(i)]) + 1];
        for (int64_t c = 0; c < ncomp; ++c) first[size_t(c) + 1] += first[size_t(c)];
        {
            std::vector<int64_t> fill(first.begin(), first.end() - 1);
            for (int64_t i = 0; i < n; ++i) members[size_t(fill[size_t(comp[size_t(i)])]++)] = i;
        }
        best.assign(size_t(ncomp), Edge{HUGE_VAL, -1, -1});
        std::atomic<int64_t> next{0};
        auto work = [&]() {
            for (int64_t c = next++; c < ncomp; c = next++) {
                auto b0 = members.begin() + first[size_t(c)], b1 = members.begin() + first[size_t(c) + 1];
                std::sort(b0, b1, [&](int64_t x, int64_t y) {
                    return lb[size_t(x)] < lb[size_t(y)] || (lb[size_t(x)] == lb[size_t(y)] && x < y);
                });
                Edge &bc = best[size_t(c)];
                for (auto it = b0; it != b1; ++it) {
                    const int64_t i = *it;
                    // strict: a point at exactly the best distance may hold a
                    // tie that wins on the indices
                    if (lb[size_t(i)] > bc.w2) break;
                    const double bound = bc.w2;
                    const Edge was = bc;
                    tree.nearest_foreign(i, c, comp, bc);
                    const bool found = bc.w2 != was.w2 || bc.a != was.a || bc.b != was.b;
                    lb[size_t(i)] = std::max(lb[size_t(i)], found ? bc.w2 : bound);
                }
            }
        };
        const unsigned nt = (n >= 20000 && ncomp >= 64) ? hw : 1u;
        if (nt > 1) {
            std::vector<std::thread> pool;
            for (unsigned t = 0; t < nt; ++t) pool.emplace_back(work);
            for (std::thread &t : pool) t.join();
        } else {
            work();
        }
        picks.clear();
        for (const Edge &e : best)
            if (e.a >= 0) picks.push_back(e);
        // several components may pick one edge, and ties could close a cycle:
        // add in order, only edges that join two different components
        std::sort(picks.begin(), picks.end(), lighter);
        const int64_t before = ne;
        for (const Edge &e : picks)
            if (sets.unite(e.a, e.b)) {
                edges[2 * ne] = e.a;
                edges[2 * ne + 1] = e.b;
                ++ne;
            }
        if (ne == before) return MH_ERR_ARG;  // non-finite coordinates
    }
    return MH_OK;
}

// float dot product, summed left to right without fused multiply-adds
inline float dot3(const float *a, const float *b)
{
#pragma clang fp contract(off)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

}  // namespace

MH_EXPORT int mh_spanning_tree_host(const double *pts, int64_t n, double r0, int64_t *edges)
{
    if (n < 0 || (n > 0 && (!pts || (n > 1 && !edges))) || !(r0 >= 0.0)) return MH_ERR_ARG;
    if (n < 2) return MH_OK;
    KdTree tree(pts, n);
    DisjointSets sets(n);
    int64_t ne = 0;
    std::vector<Edge> shorts;
    const double r = r0 * (1.0 + 1e-9);
    tree.pairs_within(r * r, shorts);
    kruskal(shorts, sets, edges, ne);
    return complete_tree(tree, n, sets, edges, ne);
}

// Lattice points (voxel coordinates, scaled by the voxel size): the short
// edges come from the 13 forward neighbour offsets looked up in a dense index
// volume. Every pair shorter than twice the smallest voxel size is a
// 26-neighbour pair, so Kruskal on the neighbour pairs below that length is
// the full Kruskal's prefix, and the tree is mh_spanning_tree_host's on the
// scaled points.
MH_EXPORT int mh_ridge_tree_host(const int64_t *coords, int64_t n, const int64_t dims[3],
                                 const double voxel_size[3], int64_t *edges)
{
    if (n < 0 || !mh::dims_ok(dims) || !voxel_size || (n > 0 && !coords) || (n > 1 && !edges))
        return MH_ERR_ARG;
    for (int k = 0; k < 3; ++k)
        if (!(voxel_size[k] > 0.0) || !std::isfinite(voxel_size[k])) return MH_ERR_ARG;
    if (n < 2) return MH_OK;
    const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2];
    std::vector<int32_t> at(size_t(d0 * d1 * d2), -1);
    std::vector<double> pts(size_t(3 * n));
    if (n >= (int64_t(1) << 31)) return MH_ERR_ARG;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c0 = coords[3 * i], c1 = coords[3 * i + 1], c2 = coords[3 * i + 2];
        if (c0 < 0 || c0 >= d0 || c1 < 0 || c1 >= d1 || c2 < 0 || c2 >= d2) return MH_ERR_ARG;
        int32_t &slot = at[size_t(c2 + d2 * (c1 + d1 * c0))];
        if (slot >= 0) return MH_ERR_ARG;  // repeated voxel
        slot = int32_t(i);
        pts[size_t(3 * i)] = double(c0) * voxel_size[0];
        pts[size_t(3 * i + 1)] = double(c1) * voxel_size[1];
        pts[size_t(3 * i + 2)] = double(c2) * voxel_size[2];
    }
    const double vmin = std::min(voxel_size[0], std::min(voxel_size[1], voxel_size[2]));
    const double limit = (2.0 * vmin) * (2.0 * vmin) * (1.0 - 1e-9);
    std::vector<Edge> shorts;
    shorts.reserve(size_t(4 * n));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c0 = coords[3 * i], c1 = coords[3 * i + 1], c2 = coords[3 * i + 2];
        for (int o0 = 0; o0 <= 1; ++o0)
            for (int o1 = -1; o1 <= 1; ++o1)
                for (int o2 = -1; o2 <= 1; ++o2) {
                    // forward offsets only: (0,0,1), (0,1,*), (1,*,*)
                    if (o0 == 0 && (o1 < 0 || (o1 == 0 && o2 <= 0))) continue;
                    const int64_t e0 = c0 + o0, e1 = c1 + o1, e2 = c2 + o2;
                    if (e0 >= d0 || e1 < 0 || e1 >= d1 || e2 < 0 || e2 >= d2) continue;
                    const int32_t j = at[size_t(e2 + d2 * (e1 + d1 * e0))];
                    if (j < 0) continue;
                    const double *a = &pts[size_t(3 * i)], *b = &pts[size_t(3 * j)];
                    const double x0 = a[0] - b[0], x1 = a[1] - b[1], x2 = a[2] - b[2];
                    const double w2 = x0 * x0 + x1 * x1 + x2 * x2;
                    if (w2 < limit) shorts.push_back(Edge{w2, std::min(i, int64_t(j)), std::max(i, int64_t(j))});
                }
    }
    std::vector<int32_t>().swap(at);
    DisjointSets sets(n);
    int64_t ne = 0;
    kruskal(shorts, sets, edges, ne);
    std::vector<Edge>().swap(shorts);
    KdTree tree(pts.data(), n);
    return complete_tree(tree, n, sets, edges, ne);
}

// Signs along the tree from `root`: a child keeps its direction when its dot
// product with its adjusted parent is >= 0 and flips otherwise (a NaN product
// flips), processed in breadth-first order from the root. Points not reached
// from the root are NaN.
MH_EXPORT int mh_tree_signs_host(const float *dirs, int64_t n, const int64_t *edges, int64_t ne,
                                 int64_t root, float *out)
{
    if (n < 0 || ne < 0 || (n > 0 && (!dirs || !out || root < 0 || root >= n)) || (ne > 0 && !edges))
        return MH_ERR_ARG;
    if (n == 0) return MH_OK;
    std::vector<int64_t> start(size_t(n) + 1, 0), adj(size_t(2 * ne));
    for (int64_t k = 0; k < 2 * ne; ++k) {
        if (edges[k] < 0 || edges[k] >= n) return MH_ERR_ARG;
        ++start[size_t(edges[k]) + 1];
    }
    for (int64_t i = 0; i < n; ++i) start[size_t(i) + 1] += start[size_t(i)];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t k = 0; k < ne; ++k) {
        const int64_t a = edges[2 * k], b = edges[2 * k + 1];
        adj[size_t(fill[size_t(a)]++)] = b;
        adj[size_t(fill[size_t(b)]++)] = a;
    }
    const float nan = std::numeric_limits<float>::quiet_NaN();
    for (int64_t i = 0; i < 3 * n; ++i) out[i] = nan;
    std::vector<char> seen(size_t(n), 0);
    std::vector<int64_t> queue;
    queue.reserve(size_t(n));
    std::memcpy(out + 3 * root, dirs + 3 * root, 3 * sizeof(float));
    seen[size_t(root)] = 1;
    queue.push_back(root);
    for (size_t h = 0; h < queue.size(); ++h) {
        const int64_t p = queue[h];
        const float *pp = out + 3 * p;
        for (int64_t t = start[size_t(p)]; t < start[size_t(p) + 1]; ++t) {
            const int64_t c = adj[size_t(t)];
            if (seen[size_t(c)]) continue;
            seen[size_t(c)] = 1;
            const float *cd = dirs + 3 * c;
            const float dot = dot3(cd, pp);
            const float s = (dot >= 0.f) ? 1.f : -1.f;
            out[3 * c] = s * cd[0];
            out[3 * c + 1] = s * cd[1];
            out[3 * c + 2] = s * cd[2];
            queue.push_back(c);
        }
    }
    return MH_OK;
}
