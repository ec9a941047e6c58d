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
    explicit DisjointSets(int64_t n) : parent(static_cast<size_t>(n))
    {
        std::iota(parent.begin(), parent.end(), 0);
    }
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

unsigned worker_count()
{
    return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// f(k) for every k in [0, count), dealt to up to `threads` threads
template <class F>
void parallel_for(int64_t count, unsigned threads, F f)
{
    if (threads <= 1 || count <= 1) {
        for (int64_t k = 0; k < count; ++k) f(k);
        return;
    }
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (int64_t k = next++; k < count; k = next++) f(k);
    };
    std::vector<std::thread> pool;
    const int64_t nt = std::min<int64_t>(threads, count);
    for (int64_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
}

// k-d tree over the points, stored contiguously in tree order. Node ids
// follow a fixed pre-order layout (left child id + 1, right child after the
// left subtree), so subtrees can be built by different threads. During a
// Boruvka round every node knows the component shared by all its points
// (-1 when mixed).
class KdTree {
public:
    static constexpr int64_t kLeaf = 8;
    struct Point {
        double x[3];
        int64_t id;
    };
    struct Node {
        double lo[3], hi[3];
        int64_t begin, end;
        int64_t left, right;  // -1 for leaves
    };

    KdTree(const double *pts, int64_t n) : p_(static_cast<size_t>(n))
    {
        for (int64_t i = 0; i < n; ++i)
            p_[size_t(i)] = Point{{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, i};
        nodes_.resize(size_t(count_nodes(n)));
        build(0, 0, n, 0);
        comp_.assign(nodes_.size(), -1);
        pcomp_.assign(p_.size(), -1);
    }

    int64_t size() const { return int64_t(p_.size()); }
    int64_t id_at(int64_t t) const { return p_[size_t(t)].id; }

    // every pair (i, j), i < j, with |pi - pj|^2 <= r2
    void pairs_within(double r2, std::vector<Edge> &out) const
    {
        for (int64_t t = 0; t < size(); ++t) pairs_of(0, t, r2, out);
    }

    // components of the points (by point id) and of every node
    void label(const std::vector<int64_t> &comp)
    {
        for (size_t t = 0; t < p_.size(); ++t) pcomp_[t] = comp[size_t(p_[t].id)];
        // children have larger ids than their parent
        for (int64_t id = int64_t(nodes_.size()) - 1; id >= 0; --id) {
            const Node &nd = nodes_[size_t(id)];
            int64_t c;
            if (nd.left < 0) {
                c = pcomp_[size_t(nd.begin)];
                for (int64_t t = nd.begin + 1; t < nd.end && c >= 0; ++t)
                    if (pcomp_[size_t(t)] != c) c = -1;
            } else {
                const int64_t a = comp_[size_t(nd.left)], b = comp_[size_t(nd.right)];
                c = (a == b) ? a : -1;
            }
            comp_[size_t(id)] = c;
        }
    }
    int64_t comp_at(int64_t t) const { return pcomp_[size_t(t)]; }

    // shortest edge from the point at tree position t to a point of another
    // component, if lighter than `best`
    void nearest_foreign(int64_t t, Edge &best) const { foreign(0, t, pcomp_[size_t(t)], best); }

private:
    std::vector<Point> p_;
    std::vector<Node> nodes_;
    std::vector<int64_t> comp_, pcomp_;

    static int64_t count_nodes(int64_t m)
    {
        if (m <= kLeaf) return 1;
        return 1 + count_nodes(m / 2) + count_nodes(m - m / 2);
    }

    void build(int64_t id, int64_t b, int64_t e, int depth)
    {
        Node nd;
        nd.begin = b;
        nd.end = e;
        nd.left = nd.right = -1;
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = HUGE_VAL;
            nd.hi[k] = -HUGE_VAL;
        }
        for (int64_t t = b; t < e; ++t)
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::min(nd.lo[k], p_[size_t(t)].x[k]);
                nd.hi[k] = std::max(nd.hi[k], p_[size_t(t)].x[k]);
            }
        if (e - b > kLeaf) {
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (nd.hi[k] - nd.lo[k] > nd.hi[axis] - nd.lo[axis]) axis = k;
            const int64_t m = b + (e - b) / 2;
            std::nth_element(p_.begin() + b, p_.begin() + m, p_.begin() + e,
                             [axis](const Point &x, const Point &y) {
                                 return x.x[axis] < y.x[axis] ||
                                        (x.x[axis] == y.x[axis] && x.id < y.id);
                             });
            nd.left = id + 1;
            nd.right = id + 1 + count_nodes(m - b);
            if (depth < 4 && e - b > (int64_t(1) << 16)) {
                const int64_t left = nd.left;
                std::thread th([this, left, b, m, depth]() { build(left, b, m, depth + 1); });
                build(nd.right, m, e, depth + 1);
                th.join();
            } else {
                build(nd.left, b, m, depth + 1);
                build(nd.right, m, e, depth + 1);
            }
        }
        nodes_[size_t(id)] = nd;
    }

    static double box_dist2(const Node &nd, const double *q)
    {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = q[k] < nd.lo[k] ? nd.lo[k] - q[k] : (q[k] > nd.hi[k] ? q[k] - nd.hi[k] : 0.0);
            s += d * d;
        }
        return s;
    }

    static double dist2(const double *a, const double *b)
    {
        const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        return d0 * d0 + d1 * d1 + d2 * d2;
    }

    void pairs_of(int64_t id, int64_t t, double r2, std::vector<Edge> &out) const
    {
        const Node &nd = nodes_[size_t(id)];
        const Point &p = p_[size_t(t)];
        if (box_dist2(nd, p.x) > r2) return;
        if (nd.left < 0) {
            for (int64_t u = nd.begin; u < nd.end; ++u) {
                const Point &q = p_[size_t(u)];
                if (q.id > p.id) {
                    const double d = dist2(p.x, q.x);
                    if (d <= r2) out.push_back(Edge{d, p.id, q.id});
                }
            }
            return;
        }
        pairs_of(nd.left, t, r2, out);
        pairs_of(nd.right, t, r2, out);
    }

    void foreign(int64_t id, int64_t t, int64_t c, Edge &best) const
    {
        if (comp_[size_t(id)] == c) return;
        const Node &nd = nodes_[size_t(id)];
        const Point &p = p_[size_t(t)];
        // strict: a box at exactly the best distance may hold a tie that
        // wins on the indices
        if (box_dist2(nd, p.x) > best.w2) return;
        if (nd.left < 0) {
            for (int64_t u = nd.begin; u < nd.end; ++u) {
                if (pcomp_[size_t(u)] == c) continue;
                const Point &q = p_[size_t(u)];
                const Edge e{dist2(p.x, q.x), std::min(p.id, q.id), std::max(p.id, q.id)};
                if (lighter(e, best)) best = e;
            }
            return;
        }
        const double dl = box_dist2(nodes_[size_t(nd.left)], p.x);
        const double dr = box_dist2(nodes_[size_t(nd.right)], p.x);
        if (dl <= dr) {
            foreign(nd.left, t, c, best);
            foreign(nd.right, t, c, best);
        } else {
            foreign(nd.right, t, c, best);
            foreign(nd.left, t, c, best);
        }
    }
};

void add_edge(const Edge &e, DisjointSets &sets, int64_t *edges, int64_t &ne)
{
    if (sets.unite(e.a, e.b)) {
        edges[2 * ne] = e.a;
        edges[2 * ne + 1] = e.b;
        ++ne;
    }
}

// Kruskal over `shorts` (sorted here), extending the forest in `sets`.
void kruskal(std::vector<Edge> &shorts, DisjointSets &sets, int64_t *edges, int64_t &ne)
{
    std::sort(shorts.begin(), shorts.end(), lighter);
    for (const Edge &e : shorts) add_edge(e, sets, edges, ne);
}

// Boruvka rounds that join the forest's components into one tree.
//
// Per point, a lower bound on its distance^2 to the nearest point of another
// component. Components only merge, so the set of foreign points only
// shrinks and a bound stays valid in later rounds. A component's points are
// searched in increasing bound order, and the search stops once the bound
// exceeds the component's best edge so far. Components are dealt to worker
// threads; each component's search is sequential, so the result does not
// depend on the thread count.
int complete_tree(KdTree &tree, int64_t n, DisjointSets &sets, int64_t *edges, int64_t ne)
{
    std::vector<int64_t> comp(static_cast<size_t>(n)), root_ix(static_cast<size_t>(n), -1);
    std::vector<int64_t> members(static_cast<size_t>(n)), first;
    std::vector<double> lb(static_cast<size_t>(n), 0.0);  // by tree position
    std::vector<Edge> best, picks;
    struct Task {
        int64_t c, begin, end;
        Edge best;
    };
    constexpr int64_t kSlice = 32768;
    std::vector<Task> tasks;
    const unsigned hw = worker_count();
    while (ne < n - 1) {
        int64_t ncomp = 0;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t r = sets.find(i);
            if (root_ix[size_t(r)] < 0) root_ix[size_t(r)] = ncomp++;
            comp[size_t(i)] = root_ix[size_t(r)];
        }
        for (int64_t i = 0; i < n; ++i) root_ix[size_t(sets.find(i))] = -1;
        tree.label(comp);
        // tree positions of each component's points, by increasing bound
        // (then position: the order only decides how soon the search stops)
        first.assign(size_t(ncomp) + 1, 0);
        for (int64_t t = 0; t < n; ++t) ++first[size_t(tree.comp_at(t)) + 1];
        for (int64_t c = 0; c < ncomp; ++c) first[size_t(c) + 1] += first[size_t(c)];
        {
            std::vector<int64_t> fill(first.begin(), first.end() - 1);
            for (int64_t t = 0; t < n; ++t) members[size_t(fill[size_t(tree.comp_at(t))]++)] = t;
        }
        // a large component is searched in slices by several threads, each
        // slice with its own best; a component's edge is the lightest of its
        // slices' (each slice finds its exact minimum below its bound)
        tasks.clear();
        for (int64_t c = 0; c < ncomp; ++c)
            for (int64_t b = first[size_t(c)]; b < first[size_t(c) + 1]; b += kSlice)
                tasks.push_back(Task{c, b, std::min(b + kSlice, first[size_t(c) + 1]),
                                     Edge{HUGE_VAL, -1, -1}});
        auto search = [&](int64_t k) {
            Task &task = tasks[size_t(k)];
            auto b0 = members.begin() + task.begin, b1 = members.begin() + task.end;
            auto by_bound = [&](int64_t x, int64_t y) {
                return lb[size_t(x)] < lb[size_t(y)] || (lb[size_t(x)] == lb[size_t(y)] && x < y);
            };
            if (!std::is_sorted(b0, b1, by_bound)) std::sort(b0, b1, by_bound);
            // a private copy: neighbouring entries belong to other threads
            Edge bc = task.best;
            for (auto it = b0; it != b1; ++it) {
                const int64_t t = *it;
                // strict: a point at exactly the best distance may hold a tie
                // that wins on the indices
                if (lb[size_t(t)] > bc.w2) break;
                const double bound = bc.w2;
                const Edge was = bc;
                tree.nearest_foreign(t, bc);
                const bool found = bc.w2 != was.w2 || bc.a != was.a || bc.b != was.b;
                lb[size_t(t)] = std::max(lb[size_t(t)], found ? bc.w2 : bound);
            }
            task.best = bc;
        };
        parallel_for(int64_t(tasks.size()), (n >= 20000 && tasks.size() >= 2) ? hw : 1u, search);
        best.assign(size_t(ncomp), Edge{HUGE_VAL, -1, -1});
        for (const Task &task : tasks)
            if (task.best.a >= 0 && lighter(task.best, best[size_t(task.c)])) best[size_t(task.c)] = task.best;
        picks.clear();
        for (const Edge &e : best)
            if (e.a >= 0) picks.push_back(e);
        // several components may pick one edge, and ties could close a cycle:
        // add in order, only edges that join two different components
        std::sort(picks.begin(), picks.end(), lighter);
        const int64_t before = ne;
        for (const Edge &e : picks) add_edge(e, sets, edges, ne);
        if (ne == before) return MH_ERR_ARG;  // non-finite coordinates
    }
    return MH_OK;
}

// The forest of the lattice's short edges: pairs of 26-neighbours shorter
// than twice the smallest voxel size (every pair that short is a 26-neighbour
// pair), added in the full Kruskal order. pts receives the scaled positions.
// With point indices in C order of the voxels, each forward offset's edges
// come out in (a, b) order, so the offsets of one exact length are merged in
// order without a sort.
int lattice_forest(const int64_t *coords, int64_t n, const int64_t dims[3], const double vs[3],
                   std::vector<double> &pts, DisjointSets &sets, int64_t *edges, int64_t &ne)
{
    const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2];
    if (n >= (int64_t(1) << 31)) return MH_ERR_ARG;
    std::vector<int32_t> at(static_cast<size_t>(d0 * d1 * d2), -1);
    pts.resize(size_t(3 * n));
    bool c_order = true;
    int64_t prev = -1;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c0 = coords[3 * i], c1 = coords[3 * i + 1], c2 = coords[3 * i + 2];
        if (c0 < 0 || c0 >= d0 || c1 < 0 || c1 >= d1 || c2 < 0 || c2 >= d2) return MH_ERR_ARG;
        const int64_t lin = c2 + d2 * (c1 + d1 * c0);
        int32_t &slot = at[size_t(lin)];
        if (slot >= 0) return MH_ERR_ARG;  // repeated voxel
        slot = int32_t(i);
        c_order = c_order && lin > prev;
        prev = lin;
        pts[size_t(3 * i)] = double(c0) * vs[0];
        pts[size_t(3 * i + 1)] = double(c1) * vs[1];
        pts[size_t(3 * i + 2)] = double(c2) * vs[2];
    }
    const double vmin = std::min(vs[0], std::min(vs[1], vs[2]));
    const double limit = (2.0 * vmin) * (2.0 * vmin) * (1.0 - 1e-9);
    // the 13 forward offsets: (0,0,1), (0,1,*), (1,*,*)
    struct Off {
        int o0, o1, o2;
    };
    std::vector<Off> offs;
    for (int o0 = 0; o0 <= 1; ++o0)
        for (int o1 = -1; o1 <= 1; ++o1)
            for (int o2 = -1; o2 <= 1; ++o2)
                if (!(o0 == 0 && (o1 < 0 || (o1 == 0 && o2 <= 0)))) offs.push_back(Off{o0, o1, o2});
    std::vector<std::vector<Edge>> lists(offs.size());
    parallel_for(int64_t(offs.size()), worker_count(), [&](int64_t k) {
        const Off &o = offs[size_t(k)];
        std::vector<Edge> &L = lists[size_t(k)];
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e0 = coords[3 * i] + o.o0, e1 = coords[3 * i + 1] + o.o1,
                          e2 = coords[3 * i + 2] + o.o2;
            if (e0 >= d0 || e1 < 0 || e1 >= d1 || e2 < 0 || e2 >= d2) continue;
            const int32_t j = at[size_t(e2 + d2 * (e1 + d1 * e0))];
            if (j < 0) continue;
            const double *a = &pts[size_t(3 * i)], *b = &pts[size_t(3 * j)];
            const double x0 = a[0] - b[0], x1 = a[1] - b[1], x2 = a[2] - b[2];
            const double w2 = x0 * x0 + x1 * x1 + x2 * x2;
            if (w2 < limit) L.push_back(Edge{w2, std::min(i, int64_t(j)), std::max(i, int64_t(j))});
        }
    });
    std::vector<int32_t>().swap(at);
    // without a sort only when every offset's edges share one exact length
    // (and C order keeps each list in (a, b) order)
    bool uniform = c_order;
    std::vector<double> len(offs.size(), -1.0);
    for (size_t k = 0; k < lists.size() && uniform; ++k)
        for (const Edge &e : lists[k]) {
            if (len[k] < 0) len[k] = e.w2;
            if (e.w2 != len[k]) {
                uniform = false;
                break;
            }
        }
    if (!uniform) {
        std::vector<Edge> all;
        for (auto &L : lists) all.insert(all.end(), L.begin(), L.end());
        kruskal(all, sets, edges, ne);
        return MH_OK;
    }
    std::vector<double> lens;
    for (double l : len)
        if (l >= 0) lens.push_back(l);
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    for (double l : lens) {
        std::vector<size_t> cls, pos;
        for (size_t k = 0; k < offs.size(); ++k)
            if (len[k] == l) cls.push_back(k);
        pos.assign(cls.size(), 0);
        while (true) {
            size_t pick = cls.size();
            for (size_t q = 0; q < cls.size(); ++q) {
                const std::vector<Edge> &L = lists[cls[q]];
                if (pos[q] >= L.size()) continue;
                if (pick == cls.size() || lighter(L[pos[q]], lists[cls[pick]][pos[pick]])) pick = q;
            }
            if (pick == cls.size()) break;
            add_edge(lists[cls[pick]][pos[pick]++], sets, edges, ne);
        }
    }
    return MH_OK;
}

// float dot product, summed left to right without fused multiply-adds
inline float dot3(const float *a, const float *b)
{
#pragma clang fp contract(off)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

bool voxel_size_ok(const double *vs)
{
    if (!vs) return false;
    for (int k = 0; k < 3; ++k)
        if (!(vs[k] > 0.0) || !std::isfinite(vs[k])) return false;
    return true;
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
// volume. Kruskal on the neighbour pairs below twice the smallest voxel size
// is the full Kruskal's prefix, so the tree is mh_spanning_tree_host's on the
// scaled points.
MH_EXPORT int mh_ridge_tree_host(const int64_t *coords, int64_t n, const int64_t dims[3],
                                 const double voxel_size[3], int64_t *edges)
{
    if (n < 0 || !mh::dims_ok(dims) || !voxel_size_ok(voxel_size) || (n > 0 && !coords) ||
        (n > 1 && !edges))
        return MH_ERR_ARG;
    if (n < 2) return MH_OK;
    std::vector<double> pts;
    DisjointSets sets(n);
    int64_t ne = 0;
    const int rc = lattice_forest(coords, n, dims, voxel_size, pts, sets, edges, ne);
    if (rc) return rc;
    if (ne == n - 1) return MH_OK;
    KdTree tree(pts.data(), n);
    return complete_tree(tree, n, sets, edges, ne);
}

MH_EXPORT int64_t mh_ridge_forest_host(const int64_t *coords, int64_t n, const int64_t dims[3],
                                       const double voxel_size[3], int64_t *edges)
{
    if (n < 0 || !mh::dims_ok(dims) || !voxel_size_ok(voxel_size) || (n > 0 && !coords) ||
        (n > 1 && !edges))
        return MH_ERR_ARG;
    if (n < 2) return 0;
    std::vector<double> pts;
    DisjointSets sets(n);
    int64_t ne = 0;
    const int rc = lattice_forest(coords, n, dims, voxel_size, pts, sets, edges, ne);
    return rc ? rc : ne;
}

// mh_ridge_tree_host with its first phase given: `forest` (nf edges) is
// mh_ridge_forest_host's output for these points, or that output for a
// superset restricted to a union of its trees and renumbered in order (the
// forest of the subset: a tree's Kruskal decisions only involve its own
// edges, and renumbering in order keeps their order). The forest is copied
// to the front of `edges` and the Boruvka rounds add the rest, so the edges
// equal mh_ridge_tree_host's. Checked: indices in range, coordinates inside
// dims, no edge closing a cycle and every edge shorter than twice the
// smallest voxel size; the points must be distinct, as there.
MH_EXPORT int mh_ridge_tree_from_forest_host(const int64_t *coords, int64_t n, const int64_t dims[3],
                                             const double voxel_size[3], const int64_t *forest,
                                             int64_t nf, int64_t *edges)
{
    if (n < 0 || nf < 0 || nf > std::max<int64_t>(n - 1, 0) || !mh::dims_ok(dims) ||
        !voxel_size_ok(voxel_size) || (n > 0 && !coords) || (n > 1 && !edges) || (nf > 0 && !forest))
        return MH_ERR_ARG;
    if (n < 2) return MH_OK;
    if (n >= (int64_t(1) << 31)) return MH_ERR_ARG;
    std::vector<double> pts(size_t(3 * n));
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const int64_t c = coords[3 * i + k];
            if (c < 0 || c >= dims[k]) return MH_ERR_ARG;
            pts[size_t(3 * i + k)] = double(c) * voxel_size[k];
        }
    const double vmin = std::min(voxel_size[0], std::min(voxel_size[1], voxel_size[2]));
    const double limit = (2.0 * vmin) * (2.0 * vmin) * (1.0 - 1e-9);
    DisjointSets sets(n);
    int64_t ne = 0;
    for (int64_t k = 0; k < nf; ++k) {
        const int64_t a = forest[2 * k], b = forest[2 * k + 1];
        if (a < 0 || a >= n || b < 0 || b >= n) return MH_ERR_ARG;
        const double *p = &pts[size_t(3 * a)], *q = &pts[size_t(3 * b)];
        const double x0 = p[0] - q[0], x1 = p[1] - q[1], x2 = p[2] - q[2];
        if (!(x0 * x0 + x1 * x1 + x2 * x2 < limit)) return MH_ERR_ARG;
        const int64_t before = ne;
        add_edge(Edge{0.0, a, b}, sets, edges, ne);
        if (ne == before) return MH_ERR_ARG;  // closes a cycle
    }
    if (ne == n - 1) return MH_OK;
    KdTree tree(pts.data(), n);
    return complete_tree(tree, n, sets, edges, ne);
}

// Disjoint sets whose links carry a sign parity: parity(x, y) = 1 when x and
// y take opposite signs. A set's root is its smallest index.
struct ParitySets {
    std::vector<int32_t> parent;
    std::vector<uint8_t> par;  // parity to the parent
    explicit ParitySets(int64_t n) : parent(static_cast<size_t>(n)), par(static_cast<size_t>(n), 0)
    {
        for (int64_t i = 0; i < n; ++i) parent[size_t(i)] = int32_t(i);
    }
    // root of x, with x's parity to it; points the path at the root
    int32_t find(int32_t x, uint8_t &px)
    {
        int32_t r = x;
        uint8_t p = 0;
        while (parent[size_t(r)] != r) {
            p ^= par[size_t(r)];
            r = parent[size_t(r)];
        }
        px = p;
        uint8_t q = p;  // parity of x to r, then of each next node on the path
        while (parent[size_t(x)] != r && parent[size_t(x)] != x) {
            const int32_t nx = parent[size_t(x)];
            const uint8_t qn = q ^ par[size_t(x)];
            parent[size_t(x)] = r;
            par[size_t(x)] = q;
            x = nx;
            q = qn;
        }
        return r;
    }
    // joins the sets of a and b so that parity(a, b) = p (no-op once joined)
    void join(int32_t a, int32_t b, uint8_t p)
    {
        uint8_t pa, pb;
        const int32_t ra = find(a, pa), rb = find(b, pb);
        if (ra == rb) return;
        const uint8_t link = p ^ pa ^ pb;  // parity(ra, rb)
        if (ra < rb) {
            parent[size_t(rb)] = ra;
            par[size_t(rb)] = link;
        } else {
            parent[size_t(ra)] = rb;
            par[size_t(ra)] = link;
        }
    }
};

// The forest's edges with a non-zero, non-NaN direction product joined in
// parity sets (negative product: opposite signs); the others returned as
// cuts, with the sign the walk below gives the child across them (zero: the
// child keeps its direction, NaN: it flips). Products are formed as the walk
// forms them (dot3).
struct Cut {
    int32_t a, b;
    int8_t sign;
};
int join_forest(const float *dirs, int64_t n, const int64_t *edges, int64_t ne, ParitySets &S,
                std::vector<Cut> &cuts)
{
    for (int64_t k = 0; k < ne; ++k) {
        const int64_t a = edges[2 * k], b = edges[2 * k + 1];
        if (a < 0 || a >= n || b < 0 || b >= n) return MH_ERR_ARG;
        const float d = dot3(dirs + 3 * a, dirs + 3 * b);
        if (d != d) cuts.push_back(Cut{int32_t(a), int32_t(b), -1});
        else if (d == 0.f) cuts.push_back(Cut{int32_t(a), int32_t(b), 1});
        else S.join(int32_t(a), int32_t(b), uint8_t(d < 0.f ? 1 : 0));
    }
    return MH_OK;
}

// Pieces of a forest between its edges whose direction product is zero or
// NaN, and signs within each piece: every edge of a piece joins points whose
// signed directions have a non-negative product, a piece's smallest point
// keeping its direction. In a forest without zero or NaN products this is
// the walk of mh_tree_signs_host from root 0, bit for bit. counts[0]: zero
// products, counts[1]: NaN products. Pieces are labelled 0 .. P-1 in order
// of their smallest point.
MH_EXPORT int mh_forest_pieces_host(const float *dirs, int64_t n, const int64_t *edges, int64_t ne,
                                    float *out, int32_t *pieces, int64_t counts[2])
{
    if (n < 0 || ne < 0 || (n > 0 && (!dirs || !out || !pieces)) || (ne > 0 && !edges) || !counts)
        return MH_ERR_ARG;
    if (n >= (int64_t(1) << 31)) return MH_ERR_ARG;
    ParitySets S(n);
    std::vector<Cut> cuts;
    const int rc = join_forest(dirs, n, edges, ne, S, cuts);
    if (rc) return rc;
    counts[0] = counts[1] = 0;
    for (const Cut &c : cuts) ++counts[c.sign > 0 ? 0 : 1];
    std::vector<int32_t> label(static_cast<size_t>(n), -1);
    int32_t next = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint8_t p;
        const int32_t r = S.find(int32_t(i), p);
        if (label[size_t(r)] < 0) label[size_t(r)] = next++;
        pieces[i] = label[size_t(r)];
        const float sg = p ? -1.f : 1.f;
        out[3 * i] = sg * dirs[3 * i];
        out[3 * i + 1] = sg * dirs[3 * i + 1];
        out[3 * i + 2] = sg * dirs[3 * i + 2];
    }
    return MH_OK;
}

// Signs along the trees of a forest, as a breadth-first walk from `root`
// and then from every point not reached yet, in index order, gives them: the
// start keeps its direction, and a child keeps its direction when its dot
// product with its adjusted parent is >= 0 and flips otherwise (a NaN
// product flips). Negating the parent negates that product exactly, so
// within a piece between zero and NaN products every sign is the parity of
// its path (parity sets), and the walk only decides each piece's anchor: the
// start, or the child end of the zero or NaN edge it is entered through
// (which keeps, or flips, its own direction). The pieces are walked instead
// of the points.
MH_EXPORT int mh_tree_signs_host(const float *dirs, int64_t n, const int64_t *edges, int64_t ne,
                                 int64_t root, float *out)
{
    if (n < 0 || ne < 0 || (n > 0 && (!dirs || !out || root < 0 || root >= n)) || (ne > 0 && !edges))
        return MH_ERR_ARG;
    if (n == 0) return MH_OK;
    if (n >= (int64_t(1) << 31)) return MH_ERR_ARG;
    ParitySets S(n);
    std::vector<Cut> cuts;
    const int rc = join_forest(dirs, n, edges, ne, S, cuts);
    if (rc) return rc;
    // the cuts as adjacency between pieces (keyed by their roots)
    std::vector<int32_t> start(static_cast<size_t>(n) + 1, 0);
    std::vector<int32_t> rootA(cuts.size()), rootB(cuts.size());
    for (size_t k = 0; k < cuts.size(); ++k) {
        uint8_t p;
        rootA[k] = S.find(cuts[k].a, p);
        rootB[k] = S.find(cuts[k].b, p);
        ++start[size_t(rootA[k]) + 1];
        ++start[size_t(rootB[k]) + 1];
    }
    for (int64_t i = 0; i < n; ++i) start[size_t(i) + 1] += start[size_t(i)];
    struct Entry {
        int32_t piece, entry;
        int8_t sign;
    };
    std::vector<Entry> adj(2 * cuts.size());
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (size_t k = 0; k < cuts.size(); ++k) {
            adj[size_t(fill[size_t(rootA[k])]++)] = Entry{rootB[k], cuts[k].b, cuts[k].sign};
            adj[size_t(fill[size_t(rootB[k])]++)] = Entry{rootA[k], cuts[k].a, cuts[k].sign};
        }
    }
    std::vector<int32_t> anchor(static_cast<size_t>(n), -1);  // by piece root
    std::vector<int8_t> asign(static_cast<size_t>(n), 0);
    std::vector<int32_t> queue;
    auto settle = [&](int32_t from) {
        uint8_t p;
        const int32_t r0 = S.find(from, p);
        anchor[size_t(r0)] = from;
        asign[size_t(r0)] = 1;
        queue.clear();
        queue.push_back(r0);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t r = queue[h];
            for (int32_t t = start[size_t(r)]; t < start[size_t(r) + 1]; ++t) {
                const Entry &e = adj[size_t(t)];
                if (anchor[size_t(e.piece)] >= 0) continue;
                anchor[size_t(e.piece)] = e.entry;
                asign[size_t(e.piece)] = e.sign;
                queue.push_back(e.piece);
            }
        }
    };
    settle(int32_t(root));
    for (int64_t i = 0; i < n; ++i) {
        uint8_t p;
        if (anchor[size_t(S.find(int32_t(i), p))] < 0) settle(int32_t(i));
    }
    for (int64_t i = 0; i < n; ++i) {
        uint8_t p, pa;
        const int32_t r = S.find(int32_t(i), p);
        (void)S.find(anchor[size_t(r)], pa);
        const float sg = ((p ^ pa) ? -1.f : 1.f) * float(asign[size_t(r)]);
        out[3 * i] = sg * dirs[3 * i];
        out[3 * i + 1] = sg * dirs[3 * i + 1];
        out[3 * i + 2] = sg * dirs[3 * i + 2];
    }
    return MH_OK;
}
