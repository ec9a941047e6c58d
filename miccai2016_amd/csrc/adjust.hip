// Direction adjustment on the device, from the ridge (after hysteresis) and
// its 26-connected segments to every voxel's signed main direction
// (processing/adjust_directions.py:213-256, :290-339, :365-417 and
// adjust_directions.cl:182-246), for an isotropic voxel size of a power of
// two. The result equals the host construction (ridge_tree.hip + the
// cleaning in adjust_directions.py) bit for bit; where that cannot be
// guaranteed the call returns MH_DEFER and changes nothing, and the caller
// takes the host construction.
//
//  1. Ridge points: the ridge voxels in C order (their rank is the point
//     index), with a voxel -> point map.
//  2. Spanning forest of the 26-neighbour pairs by Boruvka rounds. An edge's
//     key is (squared length class, smaller point, larger point) in 64 bits,
//     the order of the host's Kruskal (lengths 1, 2, 3 times the squared
//     voxel size, then the indices), so the forest is the host's: the
//     minimum spanning forest under a total order is unique.
//  3. Pieces: the forest without its zero-product edges, as lock-free
//     union-find whose links carry the sign parity (a negative product joins
//     opposite signs). Each point gets its piece's smallest point and its
//     parity to it.
//  4. Cleaning: per piece, the sum of the signed directions (float64) and the
//     point count; the host keeps the ceil(3 ln(e + S)) longest segments,
//     deciding only where the segment lengths' bounds (over the pieces'
//     unknown relative signs, widened by a rounding margin) separate them.
//  5. The tree between the K kept segments: for every pair of segments its
//     lightest point pair, (squared distance, smaller, larger index); pairs
//     with the largest segment through its exact separable distance
//     transform, the others by brute force; Kruskal over the K segments.
//  6. Signs: the walk from the smallest kept point over the pieces joined by
//     the tree's edges (host, one entry per piece), then every kept point's
//     sign on the device, and the nearest-ridge signs of all voxels
//     (mh_nonridge_signs).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

#include "mh_common.h"

namespace {

constexpr unsigned kNone = 0xffffffffu;
constexpr uint64_t kNoKey = ~uint64_t(0);
constexpr int kPtBits = 30;
constexpr uint64_t kPtMask = (uint64_t(1) << kPtBits) - 1;
constexpr int kScanItems = 4;
constexpr int kScanTile = mh::kBlock * kScanItems;
constexpr unsigned kCutCap = 1u << 16;      // zero-product forest edges kept for the walk
constexpr int64_t kMaxOthers = 1 << 16;     // kept points outside the largest kept segment
constexpr int kMaxRounds = 64;
constexpr int kMaxKept = 4096;              // kept segments (their pair table is K x K)
constexpr int kHostPrimMax = 64;            // larger segment trees by Boruvka on the device
constexpr int kDensePoints = 256;           // the host's dense spanning tree up to this many
constexpr int kJumpBatch = 4;               // pointer-doubling steps per host check

struct Grid {
    int d0, d1, d2;
};

// Relaxed atomics: the union-find words carry all the information a reader
// needs (every word names an ancestor with the parity to it, at any moment),
// and flags are read by the host after the kernel ends, so no ordering
// against other memory is required.
__device__ __forceinline__ unsigned ld_relaxed(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A flag set by many waves: one store per wave that asks, into one of 64
// slots 128 bytes apart chosen by the work-group (stores from thousands of
// waves to one address serialise); the host reads all slots.
constexpr int kFlagSlots = 64;
constexpr int kFlagStride = 32;

__device__ __forceinline__ void wave_flag(bool pred, unsigned *flags)
{
    const unsigned long long m = __ballot(pred);
    if (m && __lane_id() == unsigned(__builtin_ctzll(m)))
        st_relaxed(flags + (blockIdx.x & (kFlagSlots - 1)) * kFlagStride, 1u);
}

// float dot product, one rounding per operation, left to right
__host__ __device__ __forceinline__ float dot3(const float *a, const float *b)
{
#pragma clang fp contract(off)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// ------------------------------------------------------------ compaction --
// Flags in C order -> their ranks. Tile t covers [t*kScanTile, (t+1)*kScanTile),
// each thread kScanItems consecutive entries.

template <class T>
__global__ void __launch_bounds__(mh::kBlock)
flag_count_kernel(const T *__restrict__ flag, int64_t n, unsigned *__restrict__ counts)
{
    __shared__ unsigned part[mh::kBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    unsigned c = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n && flag[base + k] != T(0)) ++c;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < mh::kBlock / 64; ++w) t += part[w];
        counts[blockIdx.x] = t;
    }
}

// exclusive scan of counts[0 .. nb) in place by one work-group; total -> *total
__global__ void __launch_bounds__(1024)
scan_counts_kernel(unsigned *__restrict__ counts, unsigned nb, unsigned *__restrict__ total)
{
    __shared__ unsigned sums[1024];
    const unsigned per = (nb + 1023) / 1024;
    const unsigned b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    unsigned s = 0;
    for (unsigned b = b0; b < b1; ++b) s += counts[b];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (unsigned off = 1; off < 1024; off <<= 1) {
        const unsigned v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0u;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = sums[threadIdx.x] - s;  // exclusive prefix of this thread's range
    for (unsigned b = b0; b < b1; ++b) {
        const unsigned c = counts[b];
        counts[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023) *total = sums[1023];
}

// out[rank] = index of every flagged entry; with rank_of, rank_of[i] = rank
// or -1 for every entry
template <class T>
__global__ void __launch_bounds__(mh::kBlock)
compact_kernel(const T *__restrict__ flag, int64_t n, const unsigned *__restrict__ offs,
               int *__restrict__ out, int *__restrict__ rank_of, const int *__restrict__ src = nullptr)
{
    __shared__ unsigned scan[mh::kBlock];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    bool f[kScanItems];
    unsigned c = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        f[k] = base + k < n && flag[base + k] != T(0);
        c += f[k] ? 1u : 0u;
    }
    scan[threadIdx.x] = c;
    __syncthreads();
    for (unsigned off = 1; off < mh::kBlock; off <<= 1) {
        const unsigned v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0u;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned r = offs[blockIdx.x] + scan[threadIdx.x] - c;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k >= n) break;
        if (f[k]) out[r] = src ? src[base + k] : int(base + k);
        if (rank_of) rank_of[base + k] = f[k] ? int(r) : -1;
        r += f[k] ? 1u : 0u;
    }
}

// ------------------------------------------------------- Boruvka forest --

__global__ void __launch_bounds__(mh::kBlock)
bv_init_kernel(int R, int *__restrict__ comp, unsigned long long *__restrict__ best,
               unsigned long long *__restrict__ fkey)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    comp[p] = p;
    best[p] = kNoKey;
    fkey[p] = kNoKey;
}

// every active point's lightest pair with a 26-neighbour of another
// component, minimised per component; still[i] says whether active point i
// had one (a point whose neighbours all share its component never has one
// again: components only merge). act: the active points (nullptr: all).
__global__ void __launch_bounds__(mh::kBlock)
bv_min_edge_kernel(const int *__restrict__ act_list, int nact, const int *__restrict__ pts,
                   const int *__restrict__ at, const int *__restrict__ comp, Grid g,
                   unsigned long long *__restrict__ best, unsigned char *__restrict__ still)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    uint64_t m = kNoKey;
    const bool act = i < nact;
    const int p = act ? (act_list ? act_list[i] : i) : 0;
    const int c = act ? comp[p] : 0;
    const int v = act ? pts[p] : 0;
    const int i2 = v % g.d2, i1 = (v / g.d2) % g.d1, i0 = v / (g.d2 * g.d1);
    for (int a = act ? -1 : 2; a <= 1; ++a) {
        const int j0 = i0 + a;
        if (j0 < 0 || j0 >= g.d0) continue;
        for (int b = -1; b <= 1; ++b) {
            const int j1 = i1 + b;
            if (j1 < 0 || j1 >= g.d1) continue;
            for (int e = -1; e <= 1; ++e) {
                const int j2 = i2 + e;
                if ((a | b | e) == 0 || j2 < 0 || j2 >= g.d2) continue;
                const int q = at[(j0 * g.d1 + j1) * g.d2 + j2];
                if (q < 0 || comp[q] == c) continue;
                const uint64_t cls = uint64_t(abs(a) + abs(b) + abs(e) - 1);
                const uint64_t lo = uint64_t(min(p, q)), hi = uint64_t(max(p, q));
                const uint64_t key = (cls << 60) | (lo << kPtBits) | hi;
                m = key < m ? key : m;
            }
        }
    }
    if (m != kNoKey) atomicMin(&best[c], (unsigned long long)m);
    if (act) still[i] = m != kNoKey ? 1 : 0;
}

// each root hooks to the component at the other end of its lightest edge;
// of two roots that chose the same edge the larger hooks
__global__ void __launch_bounds__(mh::kBlock)
bv_hook_kernel(const int *__restrict__ comp, const unsigned long long *__restrict__ best, int R,
               int *__restrict__ up, unsigned long long *__restrict__ fkey)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R || comp[p] != p) return;
    const uint64_t k = best[p];
    up[p] = p;
    if (k == kNoKey) return;
    const int a = int((k >> kPtBits) & kPtMask), b = int(k & kPtMask);
    const int ca = comp[a], cb = comp[b];
    const int other = ca == p ? cb : ca;
    if (best[other] == k && p < other) return;  // the other root hooks to this one
    up[p] = other;
    fkey[p] = k;
}

// pointer doubling over the roots' hook trees (double-buffered)
__global__ void __launch_bounds__(mh::kBlock)
bv_jump_kernel(const int *__restrict__ comp, const int *__restrict__ up_in, int *__restrict__ up_out,
               int R, unsigned *__restrict__ changed)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    bool moved = false;
    if (p < R && comp[p] == p) {
        const int u = up_in[p];
        const int uu = up_in[u];
        up_out[p] = uu;
        moved = uu != u;
    }
    wave_flag(moved, changed);
}

__global__ void __launch_bounds__(mh::kBlock)
bv_relabel_kernel(int *__restrict__ comp, const int *__restrict__ up, int R,
                  unsigned long long *__restrict__ best)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    comp[p] = up[comp[p]];
    best[p] = kNoKey;
}

// ------------------------------------------------- pieces with parities --
// pw[p] = parent << 1 | parity(p, parent); a root holds itself with parity 0.

// Links only ever point at a smaller index (a root joins the smaller root,
// halving points at an ancestor), so every walk ends within R steps.
template <bool HALVE>
__device__ unsigned pw_find(unsigned *pw, unsigned x, unsigned &par, int R, unsigned *err)
{
    unsigned acc = 0;
    for (int it = 0; it <= R; ++it) {
        const unsigned w = ld_relaxed(pw + x);
        const unsigned px = w >> 1;
        if (px == x) {
            par = acc;
            return x;
        }
        const unsigned wp = ld_relaxed(pw + px);
        const unsigned gx = wp >> 1;
        if (HALVE && gx != px) {
            // halve the path: x now points at its grandparent (an ancestor,
            // with the parity of the two links; roots are never written here)
            const unsigned nw = (gx << 1) | ((w ^ wp) & 1u);
            st_relaxed(pw + x, nw);
            acc ^= nw & 1u;
            x = gx;
        } else {
            acc ^= w & 1u;
            x = px;
        }
    }
    st_relaxed(err, 1u);
    par = acc;
    return x;
}

__global__ void __launch_bounds__(mh::kBlock)
pu_init_kernel(unsigned *__restrict__ pw, int R)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < R) pw[p] = unsigned(p) << 1;
}

// forest edges: zero products are cuts (kept for the walk), NaN products
// counted, the others join their ends' pieces with the product's parity
__global__ void __launch_bounds__(mh::kBlock)
pu_union_kernel(const unsigned long long *__restrict__ fkey, int R, const int *__restrict__ pts,
                const float *__restrict__ dirs, unsigned *pw, unsigned *__restrict__ counts,
                unsigned long long *__restrict__ cuts, unsigned *__restrict__ err)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    const uint64_t k = fkey[p];
    if (k == kNoKey) return;
    const unsigned a = unsigned((k >> kPtBits) & kPtMask), b = unsigned(k & kPtMask);
    const float d = dot3(dirs + 3 * int64_t(pts[a]), dirs + 3 * int64_t(pts[b]));
    if (d != d) {
        atomicAdd(&counts[1], 1u);
        return;
    }
    if (d == 0.f) {
        const unsigned slot = atomicAdd(&counts[0], 1u);
        if (slot < kCutCap) cuts[slot] = (unsigned long long)(k & ((uint64_t(1) << 60) - 1));
        return;
    }
    const unsigned rel = d < 0.f ? 1u : 0u;
    for (int it = 0; it <= R; ++it) {
        unsigned pa, pb;
        const unsigned ra = pw_find<true>(pw, a, pa, R, err), rb = pw_find<true>(pw, b, pb, R, err);
        if (ra == rb) return;
        const unsigned hi = max(ra, rb), lo = min(ra, rb);
        const unsigned want = (lo << 1) | (pa ^ pb ^ rel);
        if (atomicCAS(pw + hi, hi << 1, want) == (hi << 1)) return;
    }
    st_relaxed(err, 1u);
}

// every point linked straight to its root. The first pass (HALVE) shortens
// paths but another thread's halving may still replace a finished entry by
// an ancestor; the second only reads others' entries, so each entry ends as
// its own thread wrote it.
template <bool HALVE>
__global__ void __launch_bounds__(mh::kBlock)
pu_flatten_kernel(unsigned *pw, int R, unsigned *__restrict__ err)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    unsigned par;
    const unsigned r = pw_find<HALVE>(pw, unsigned(p), par, R, err);
    st_relaxed(pw + p, (r << 1) | par);
}

// per piece: the sum of its points' signed directions (float64) and its
// point count; each thread sums a run of kSumRun points, flushing on a change
// of piece. The device's own float64 atomic add (the memory is the device's,
// coarse-grained), not a compare-and-swap loop.
constexpr int kSumRun = 16;

__device__ __forceinline__ void flush_piece(double *psum, unsigned *pcnt, int piece, double s0,
                                            double s1, double s2, unsigned n)
{
    unsafeAtomicAdd(psum + 3 * int64_t(piece), s0);
    unsafeAtomicAdd(psum + 3 * int64_t(piece) + 1, s1);
    unsafeAtomicAdd(psum + 3 * int64_t(piece) + 2, s2);
    atomicAdd(pcnt + piece, n);
}
__global__ void __launch_bounds__(mh::kBlock)
piece_sums_kernel(const unsigned *__restrict__ pw, const int *__restrict__ pts,
                  const float *__restrict__ dirs, int R, double *__restrict__ psum,
                  unsigned *__restrict__ pcnt)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t begin = t * kSumRun, end = min(int64_t(R), begin + kSumRun);
    int cur = -1;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    unsigned cnt = 0;
    for (int64_t p = begin; p < end; ++p) {
        const unsigned w = pw[p];
        const int r = int(w >> 1);
        if (r != cur) {
            if (cur >= 0) flush_piece(psum, pcnt, cur, s0, s1, s2, cnt);
            cur = r;
            s0 = s1 = s2 = 0.0;
            cnt = 0;
        }
        const float *d = dirs + 3 * int64_t(pts[p]);
        const float sg = (w & 1u) ? -1.f : 1.f;
        s0 += double(sg * d[0]);
        s1 += double(sg * d[1]);
        s2 += double(sg * d[2]);
        ++cnt;
    }
    // the last runs: for each distinct piece among the wave's lanes, the
    // lanes ending in it add their sums by shuffles and one lane flushes (a
    // long piece spans most lanes of most waves)
    bool pending = cur >= 0;
    for (int guard = 0; guard < 64; ++guard) {
        const unsigned long long m = __ballot(pending);
        if (!m) break;
        const int leader = __builtin_ctzll(m);
        const int key = __shfl(cur, leader);
        const bool mine = pending && cur == key;
        double a0 = mine ? s0 : 0.0, a1 = mine ? s1 : 0.0, a2 = mine ? s2 : 0.0;
        unsigned an = mine ? cnt : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            a0 += __shfl_xor(a0, off);
            a1 += __shfl_xor(a1, off);
            a2 += __shfl_xor(a2, off);
            an += __shfl_xor(an, off);
        }
        if (int(__lane_id()) == leader) flush_piece(psum, pcnt, key, a0, a1, a2, an);
        pending = pending && !mine;
    }
}

__global__ void __launch_bounds__(mh::kBlock)
piece_roots_kernel(const unsigned *__restrict__ pw, int R, int *__restrict__ roots,
                   unsigned *__restrict__ nroots)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R || (pw[p] >> 1) != unsigned(p)) return;
    roots[atomicAdd(nroots, 1u)] = p;
}

// per listed point: (pw, segment label, direction)
__global__ void __launch_bounds__(mh::kBlock)
gather_points_kernel(const int *__restrict__ list, int m, const unsigned *__restrict__ pw,
                     const int *__restrict__ pts, const int *__restrict__ labels,
                     const float *__restrict__ dirs, unsigned *__restrict__ out_pw,
                     int *__restrict__ out_lab, float *__restrict__ out_dir)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= m) return;
    const int p = list[k];
    const int v = pts[p];
    out_lab[k] = labels[v];
    if (out_pw) out_pw[k] = pw[p];
    if (out_dir) {
        out_dir[3 * k] = dirs[3 * int64_t(v)];
        out_dir[3 * k + 1] = dirs[3 * int64_t(v) + 1];
        out_dir[3 * k + 2] = dirs[3 * int64_t(v) + 2];
    }
}

__global__ void __launch_bounds__(mh::kBlock)
gather_sums_kernel(const int *__restrict__ roots, int m, const double *__restrict__ psum,
                   const unsigned *__restrict__ pcnt, double *__restrict__ out_sum,
                   unsigned *__restrict__ out_cnt)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= m) return;
    const int r = roots[k];
    out_sum[3 * k] = psum[3 * int64_t(r)];
    out_sum[3 * k + 1] = psum[3 * int64_t(r) + 1];
    out_sum[3 * k + 2] = psum[3 * int64_t(r) + 2];
    out_cnt[k] = pcnt[r];
}

// ------------------------------------------------ tree between segments --
// seg_of_root[p]: the kept-segment number (0 .. K-1) at each segment's
// smallest point, -1 elsewhere; a point's segment is found through its
// voxel's label (the segment's smallest voxel).

__device__ __forceinline__ int kept_seg(int p, const int *pts, const int *labels, const int *at,
                                        const int *seg_of_root)
{
    // labels from another ridge than `ridge` would point off it: no segment
    const int lab = labels[pts[p]];
    const int r = lab >= 0 ? at[lab] : -1;
    return r >= 0 ? seg_of_root[r] : -1;
}

// the largest kept segment's points marked with their indices; the other
// kept points bucketed by segment
__global__ void __launch_bounds__(mh::kBlock)
split_kept_kernel(const int *__restrict__ pts, const int *__restrict__ labels,
                  const int *__restrict__ at, const int *__restrict__ seg_of_root, int R, int big,
                  unsigned *__restrict__ ridx, unsigned *__restrict__ fill,
                  const unsigned *__restrict__ start, int *__restrict__ others,
                  unsigned *__restrict__ err)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    const int s = kept_seg(p, pts, labels, at, seg_of_root);
    if (s < 0) return;
    if (s == big) {
        ridx[pts[p]] = unsigned(p);
    } else {
        const unsigned slot = start[s] + atomicAdd(fill + s, 1u);
        if (slot < start[s + 1]) others[slot] = p;
        else st_relaxed(err, 1u);
    }
}

// sides below 2^14: the sum stays below 2^32
__device__ __forceinline__ unsigned sq_dist(int u, int v, Grid g)
{
    const unsigned a = unsigned(abs(u / (g.d1 * g.d2) - v / (g.d1 * g.d2)));
    const unsigned b = unsigned(abs((u / g.d2) % g.d1 - (v / g.d2) % g.d1));
    const unsigned c = unsigned(abs(u % g.d2 - v % g.d2));
    return a * a + b * b + c * c;
}

// PASS 0: the smallest squared distance of every pair of kept segments;
// PASS 1: among the point pairs at that distance, the smallest
// (smaller index, larger index). Thread (i, t): other point i against
// segment t (the largest through its transform, others when t > own).
template <int PASS>
__global__ void __launch_bounds__(mh::kBlock)
pair_kernel(const int *__restrict__ others, int m, const int *__restrict__ pts,
            const int *__restrict__ labels, const int *__restrict__ at,
            const int *__restrict__ seg_of_root, const unsigned *__restrict__ start, int K, int big,
            const unsigned *__restrict__ edt_d, const unsigned *__restrict__ edt_b, Grid g,
            unsigned *__restrict__ w2min, unsigned long long *__restrict__ idxmin)
{
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= int64_t(m) * K) return;
    const int i = int(tid / K), t = int(tid % K);
    const int p = others[i];
    const int s = kept_seg(p, pts, labels, at, seg_of_root);
    if (t == s || (t != big && t < s)) return;
    const int lo_s = min(s, t), hi_s = max(s, t);
    const int slot = lo_s * K + hi_s;
    if (t == big) {
        const int v = pts[p];
        const unsigned d = edt_d[v], q = edt_b[v];
        if (q == kNone) return;
        if (PASS == 0) {
            atomicMin(w2min + slot, d);
        } else if (d == w2min[slot]) {
            const uint64_t a = uint64_t(min(unsigned(p), q)), b = uint64_t(max(unsigned(p), q));
            atomicMin(idxmin + slot, (unsigned long long)((a << kPtBits) | b));
        }
        return;
    }
    const int v = pts[p];
    unsigned best = kNone;
    uint64_t bk = kNoKey;
    const unsigned target = PASS == 0 ? kNone : w2min[slot];
    for (unsigned k = start[t]; k < start[t + 1]; ++k) {
        const int q = others[k];
        const unsigned d = sq_dist(v, pts[q], g);
        if (PASS == 0) {
            best = d < best ? d : best;
        } else if (d == target) {
            const uint64_t a = uint64_t(min(p, q)), b = uint64_t(max(p, q));
            const uint64_t key = (a << kPtBits) | b;
            bk = key < bk ? key : bk;
        }
    }
    if (PASS == 0 && best != kNone) atomicMin(w2min + slot, best);
    if (PASS == 1 && bk != kNoKey) atomicMin(idxmin + slot, (unsigned long long)bk);
}

// ---- the segments' minimum spanning tree by Boruvka rounds on the table --
// Entry (a, b), a < b, of the K x K tables holds segments a and b's lightest
// point pair: squared distance w2min and pair idxmin. Keys compare as
// (w2, pair); per round every component takes its lightest pair to another
// component in three passes (smallest w2, smallest pair at it, its entry).

__global__ void __launch_bounds__(mh::kBlock)
sm_w2_kernel(const unsigned *__restrict__ w2min, int K, const int *__restrict__ scomp,
             unsigned *__restrict__ cw)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(K) * K) return;
    const int a = int(t / K), b = int(t % K);
    if (a >= b) return;
    const int ca = scomp[a], cb = scomp[b];
    const unsigned w = w2min[t];
    if (ca == cb || w == kNone) return;
    atomicMin(cw + ca, w);
    atomicMin(cw + cb, w);
}

__global__ void __launch_bounds__(mh::kBlock)
sm_pair_kernel(const unsigned *__restrict__ w2min, const unsigned long long *__restrict__ idxmin,
               int K, const int *__restrict__ scomp, const unsigned *__restrict__ cw,
               unsigned long long *__restrict__ cx)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(K) * K) return;
    const int a = int(t / K), b = int(t % K);
    if (a >= b) return;
    const int ca = scomp[a], cb = scomp[b];
    const unsigned w = w2min[t];
    if (ca == cb || w == kNone) return;
    if (w == cw[ca]) atomicMin(cx + ca, idxmin[t]);
    if (w == cw[cb]) atomicMin(cx + cb, idxmin[t]);
}

__global__ void __launch_bounds__(mh::kBlock)
sm_entry_kernel(const unsigned *__restrict__ w2min, const unsigned long long *__restrict__ idxmin,
                int K, const int *__restrict__ scomp, const unsigned *__restrict__ cw,
                const unsigned long long *__restrict__ cx, int64_t *__restrict__ centry)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(K) * K) return;
    const int a = int(t / K), b = int(t % K);
    if (a >= b) return;
    const int ca = scomp[a], cb = scomp[b];
    const unsigned w = w2min[t];
    if (ca == cb || w == kNone) return;
    const unsigned long long x = idxmin[t];
    if (w == cw[ca] && x == cx[ca]) centry[ca] = t;  // the keys are distinct: one entry
    if (w == cw[cb] && x == cx[cb]) centry[cb] = t;
}

// each root hooks to the component at the other end of its lightest pair;
// of two roots that chose the same pair the larger hooks
__global__ void __launch_bounds__(mh::kBlock)
sm_hook_kernel(int K, const int *__restrict__ scomp, const unsigned *__restrict__ cw,
               const unsigned long long *__restrict__ cx, const int64_t *__restrict__ centry,
               int *__restrict__ up, unsigned long long *__restrict__ edge, unsigned *__restrict__ flags)
{
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    bool hooked = false;
    if (c < K && scomp[c] == c) {
        up[c] = c;
        if (cw[c] != kNone) {
            const int64_t t = centry[c];
            const int a = int(t / K), b = int(t % K);
            const int other = scomp[a] == c ? scomp[b] : scomp[a];
            if (!(cw[other] == cw[c] && cx[other] == cx[c] && c < other)) {
                up[c] = other;
                edge[c] = cx[c];
                hooked = true;
            }
        }
    }
    wave_flag(hooked, flags);
}

__global__ void __launch_bounds__(mh::kBlock)
sm_relabel_kernel(int *__restrict__ scomp, const int *__restrict__ up, int K, unsigned *__restrict__ cw,
                  unsigned long long *__restrict__ cx)
{
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= K) return;
    scomp[c] = up[scomp[c]];
    cw[c] = kNone;
    cx[c] = kNoKey;
}

__global__ void __launch_bounds__(mh::kBlock)
sm_init_kernel(int K, int *__restrict__ scomp, unsigned *__restrict__ cw,
               unsigned long long *__restrict__ cx, unsigned long long *__restrict__ edge)
{
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= K) return;
    scomp[c] = c;
    cw[c] = kNone;
    cx[c] = kNoKey;
    edge[c] = kNoKey;
}

__global__ void __launch_bounds__(mh::kBlock)
gather_int_kernel(const int *__restrict__ idx, int m, const int *__restrict__ src, int *__restrict__ out)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < m) out[k] = src[idx[k]];
}

__global__ void __launch_bounds__(mh::kBlock)
scatter_int_kernel(const int *__restrict__ idx, const int *__restrict__ val, int m, int *__restrict__ out)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < m) out[idx[k]] = val[k];
}

// ------------------------------------------------------------ the signs --

__global__ void __launch_bounds__(mh::kBlock)
scatter_sign_kernel(const int *__restrict__ roots, const float *__restrict__ sg, int m,
                    float *__restrict__ fsign)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < m) fsign[roots[k]] = sg[k];
}

// kept flag per point (for the compaction) and the kept ridge per voxel
__global__ void __launch_bounds__(mh::kBlock)
kept_flags_kernel(const int *__restrict__ pts, const int *__restrict__ labels,
                  const int *__restrict__ at, const int *__restrict__ seg_of_root, int R,
                  unsigned char *__restrict__ kflag, float *__restrict__ kept_vox)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= R) return;
    const bool k = kept_seg(p, pts, labels, at, seg_of_root) >= 0;
    kflag[p] = k ? 1 : 0;
    if (k) kept_vox[pts[p]] = 1.f;
}

// kept points in order: coordinates (uint16) and signed directions
__global__ void __launch_bounds__(mh::kBlock)
kept_out_kernel(const int *__restrict__ kept, int m, const int *__restrict__ pts,
                const unsigned *__restrict__ pw, const float *__restrict__ fsign,
                const float *__restrict__ dirs, Grid g, uint16_t *__restrict__ coords,
                float *__restrict__ rdirs)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= m) return;
    const int p = kept[k];
    const int v = pts[p];
    const unsigned w = pw[p];
    const float sg = fsign[w >> 1] * ((w & 1u) ? -1.f : 1.f);
    coords[3 * k] = uint16_t(v / (g.d1 * g.d2));
    coords[3 * k + 1] = uint16_t((v / g.d2) % g.d1);
    coords[3 * k + 2] = uint16_t(v % g.d2);
    rdirs[3 * k] = sg * dirs[3 * int64_t(v)];
    rdirs[3 * k + 1] = sg * dirs[3 * int64_t(v) + 1];
    rdirs[3 * k + 2] = sg * dirs[3 * int64_t(v) + 2];
}

template <class T>
__global__ void __launch_bounds__(mh::kBlock) fill_kernel(T *__restrict__ x, int64_t n, T v)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// ------------------------------------------------------------- host side --

int64_t al(int64_t b) { return (b + 255) & ~int64_t(255); }

// a bump allocator over one device allocation
struct Arena {
    char *base = nullptr;
    int64_t used = 0, cap = 0;
    template <class T>
    T *take(int64_t count)
    {
        T *p = reinterpret_cast<T *>(base + used);
        used += al(int64_t(sizeof(T)) * std::max<int64_t>(count, 1));
        return used <= cap ? p : nullptr;
    }
};

// Scratch from the device's stream-ordered pool, released in stream order:
// no device-wide synchronisation (hipFree's), so the preparation workers'
// and the serving loop's other streams keep running.
struct DevMem {
    void *p = nullptr;
    hipStream_t s = nullptr;
    hipError_t alloc(int64_t bytes, hipStream_t stream)
    {
        s = stream;
        return hipMallocAsync(&p, size_t(bytes), stream);
    }
    ~DevMem()
    {
        if (p) (void)hipFreeAsync(p, s);
    }
};

struct HostSets {  // union-find with parities (host, small)
    std::vector<int> parent;
    std::vector<unsigned char> par;
    explicit HostSets(int n) : parent(size_t(n)), par(static_cast<size_t>(n), 0)
    {
        std::iota(parent.begin(), parent.end(), 0);
    }
    int find(int x, unsigned char &p)
    {
        unsigned char acc = 0;
        while (parent[size_t(x)] != x) {
            acc ^= par[size_t(x)];
            x = parent[size_t(x)];
        }
        p = acc;
        return x;
    }
    bool join(int a, int b, unsigned char rel)
    {
        unsigned char pa, pb;
        const int ra = find(a, pa), rb = find(b, pb);
        if (ra == rb) return false;
        const int hi = std::max(ra, rb), lo = std::min(ra, rb);
        parent[size_t(hi)] = lo;
        par[size_t(hi)] = pa ^ pb ^ rel;
        return true;
    }
};

bool pow2_isotropic(const double *vs)
{
    if (!vs || !(vs[0] > 0.0) || vs[0] != vs[1] || vs[1] != vs[2] || !std::isfinite(vs[0]))
        return false;
    int e;
    return std::frexp(vs[0], &e) == 0.5;
}

}  // namespace

MH_EXPORT int mh_adjust_ridge_device(mh_ctx *ctx, float *dirs, const float *ridge,
                                     const int *labels, const int64_t dims[3],
                                     const double voxel_size[3], int clean, float *kept_out,
                                     int64_t *stats, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, dirs && ridge && labels && kept_out && stats, "null pointer");
    for (int k = 0; k < 8; ++k) stats[k] = 0;
    const auto t_start = std::chrono::steady_clock::now();
    auto usec_since = [](std::chrono::steady_clock::time_point t) {
        return int64_t(std::chrono::duration_cast<std::chrono::microseconds>(
                           std::chrono::steady_clock::now() - t).count());
    };
    if (!pow2_isotropic(voxel_size)) return MH_DEFER;
    const int64_t n = dims[0] * dims[1] * dims[2];
    if (n >= (int64_t(1) << kPtBits) || dims[0] > 16384 || dims[1] > 16384 || dims[2] > 16384)
        return MH_DEFER;
    hipStream_t s = (hipStream_t)stream;
    const Grid g{int(dims[0]), int(dims[1]), int(dims[2])};
    const unsigned ntile = unsigned((n + kScanTile - 1) / kScanTile);

    // 1. ridge points
    DevMem m0;
    const int64_t bytes0 = al(4 * n) + al(4 * int64_t(ntile)) + al(4 * 8) + al(4 * n);
    MH_CHECK(ctx, m0.alloc(bytes0, s));
    Arena A0{static_cast<char *>(m0.p), 0, bytes0};
    int *at = A0.take<int>(n);
    unsigned *tcount = A0.take<unsigned>(ntile);
    unsigned *small = A0.take<unsigned>(8);
    int *pts_all = A0.take<int>(n);  // ridge points (R <= n)
    hipLaunchKernelGGL(flag_count_kernel<float>, dim3(ntile), dim3(mh::kBlock), 0, s, ridge, n, tcount);
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, tcount, ntile, small);
    hipLaunchKernelGGL(compact_kernel<float>, dim3(ntile), dim3(mh::kBlock), 0, s, ridge, n, tcount,
                       pts_all, at);
    MH_CHECK_LAUNCH(ctx, "ridge points");
    unsigned R32 = 0;
    MH_CHECK(ctx, hipMemcpyAsync(&R32, small, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    const int R = int(R32);
    stats[0] = R;
    if (R <= kDensePoints) return MH_DEFER;

    // per-point arrays
    DevMem m1;
    const unsigned rtile = unsigned((int64_t(R) + kScanTile - 1) / kScanTile);
    const int64_t bytes1 = 4 * al(4 * int64_t(R)) + 2 * al(8 * int64_t(R)) + al(4 * int64_t(R)) +
                           al(24 * int64_t(R)) + 2 * al(4 * int64_t(R)) + al(8 * int64_t(kCutCap)) +
                           al(4 * 8) + al(4 * int64_t(R)) + al(int64_t(R)) + 2 * al(4 * int64_t(R)) +
                           al(4 * kFlagSlots * kFlagStride) + al(4 * int64_t(rtile) + 4 * 8);
    MH_CHECK(ctx, m1.alloc(bytes1, s));
    Arena A1{static_cast<char *>(m1.p), 0, bytes1};
    int *comp = A1.take<int>(R);
    int *up = A1.take<int>(R);
    int *up2 = A1.take<int>(R);
    unsigned *pw = A1.take<unsigned>(R);
    auto *best = A1.take<unsigned long long>(R);
    auto *fkey = A1.take<unsigned long long>(R);
    int *roots = A1.take<int>(R);
    double *psum = A1.take<double>(3 * int64_t(R));
    unsigned *pcnt = A1.take<unsigned>(R);
    int *seg_of_root = A1.take<int>(R);
    auto *cuts = A1.take<unsigned long long>(kCutCap);
    unsigned *cnt = A1.take<unsigned>(8);  // [0] zero, [1] NaN, [2] found, [3] changed, [4] err, [5] roots
    float *fsign = A1.take<float>(R);
    unsigned char *kflag = A1.take<unsigned char>(R);
    int *actA = A1.take<int>(R);
    int *actB = A1.take<int>(R);
    unsigned *flags = A1.take<unsigned>(kFlagSlots * kFlagStride);
    unsigned *rcount = A1.take<unsigned>(int64_t(rtile) + 8);
    if (!rcount) return mh::set_error(ctx, MH_ERR_ARG, __func__, "workspace layout");
    auto any_flag = [&](bool &out) -> int {
        unsigned h[kFlagSlots * kFlagStride];
        MH_CHECK(ctx, hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
        out = false;
        for (int k = 0; k < kFlagSlots; ++k) out = out || h[k * kFlagStride] != 0;
        return MH_OK;
    };
    const int *pts = pts_all;
    const unsigned nbR = mh::blocks_for(R);
    // Pointer doubling over hook trees until no pointer moves (at most
    // log2(m) + 1 steps), in batches of kJumpBatch steps per host check: only
    // the batch's last step raises the flag, and once a step moves nothing no
    // later one does. `up` ends as the settled pointers.
    auto jump_until_settled = [&](const int *cmp, int *&up_a, int *&up_b, int m, unsigned nb) -> int {
        for (int step = 0;;) {
            if (step > 40) return mh::set_error(ctx, MH_ERR_HIP, __func__, "hook doubling");
            for (int k = 0; k < kJumpBatch; ++k, ++step) {
                if (k == kJumpBatch - 1)
                    MH_CHECK(ctx, hipMemsetAsync(flags, 0, 4 * kFlagSlots * kFlagStride, s));
                hipLaunchKernelGGL(bv_jump_kernel, dim3(nb), dim3(mh::kBlock), 0, s, cmp, up_a, up_b, m, flags);
                std::swap(up_a, up_b);
            }
            bool changed = false;
            const int rc = any_flag(changed);
            if (rc != MH_OK) return rc;
            if (!changed) return MH_OK;
        }
    };

    // 2. Boruvka rounds
    MH_CHECK(ctx, hipMemsetAsync(cnt, 0, 8 * sizeof(unsigned), s));
    hipLaunchKernelGGL(bv_init_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, R, comp, best, fkey);
    int rounds = 0;
    const int *cur = nullptr;  // the active points: all of them in the first round
    int ncur = R;
    for (;; ++rounds) {
        if (rounds >= kMaxRounds) return mh::set_error(ctx, MH_ERR_HIP, __func__, "forest rounds");
        hipLaunchKernelGGL(bv_min_edge_kernel, dim3(mh::blocks_for(ncur)), dim3(mh::kBlock), 0, s, cur, ncur,
                           pts, at, comp, g, best, kflag);
        // the points that still have an outgoing pair stay active
        int *nxt = cur == actA ? actB : actA;
        const unsigned ctile = unsigned((int64_t(ncur) + kScanTile - 1) / kScanTile);
        hipLaunchKernelGGL(flag_count_kernel<unsigned char>, dim3(ctile), dim3(mh::kBlock), 0, s, kflag,
                           int64_t(ncur), rcount);
        hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, rcount, ctile, rcount + rtile);
        hipLaunchKernelGGL(compact_kernel<unsigned char>, dim3(ctile), dim3(mh::kBlock), 0, s, kflag,
                           int64_t(ncur), rcount, nxt, (int *)nullptr, cur);
        MH_CHECK_LAUNCH(ctx, "bv_min_edge_kernel");
        unsigned nnext = 0;
        MH_CHECK(ctx, hipMemcpyAsync(&nnext, rcount + rtile, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
        if (nnext == 0) break;
        cur = nxt;
        ncur = int(nnext);
        hipLaunchKernelGGL(bv_hook_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, comp, best, R, up, fkey);
        // doubling until no pointer moves (at most log2(R) + 1 steps)
        const int rc = jump_until_settled(comp, up, up2, R, nbR);
        if (rc != MH_OK) return rc;
        hipLaunchKernelGGL(bv_relabel_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, comp, up, R, best);
        MH_CHECK_LAUNCH(ctx, "forest rounds");
    }
    stats[7] = rounds;

    // 3. pieces
    hipLaunchKernelGGL(pu_init_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, pw, R);
    hipLaunchKernelGGL(pu_union_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, fkey, R, pts, dirs, pw,
                       cnt, cuts, cnt + 4);
    hipLaunchKernelGGL(pu_flatten_kernel<true>, dim3(nbR), dim3(mh::kBlock), 0, s, pw, R, cnt + 4);
    hipLaunchKernelGGL(pu_flatten_kernel<false>, dim3(nbR), dim3(mh::kBlock), 0, s, pw, R, cnt + 4);
    // 4. sums per piece
    hipLaunchKernelGGL(fill_kernel<double>, dim3(mh::blocks_for(3 * int64_t(R))), dim3(mh::kBlock), 0, s,
                       psum, 3 * int64_t(R), 0.0);
    hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(nbR), dim3(mh::kBlock), 0, s, pcnt, int64_t(R), 0u);
    hipLaunchKernelGGL(piece_sums_kernel, dim3(mh::blocks_for((int64_t(R) + kSumRun - 1) / kSumRun)),
                       dim3(mh::kBlock),
                       0, s, pw, pts, dirs, R, psum, pcnt);
    hipLaunchKernelGGL(piece_roots_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, pw, R, roots, cnt + 5);
    MH_CHECK_LAUNCH(ctx, "pieces");
    unsigned hc[8];
    MH_CHECK(ctx, hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    if (hc[4]) return mh::set_error(ctx, MH_ERR_HIP, __func__, "union-find did not settle");
    const unsigned nzero = hc[0], nnan = hc[1];
    const int P = int(hc[5]);
    stats[4] = nzero;
    if (nnan || nzero > kCutCap) return MH_DEFER;

    // pieces to the host: roots (sorted), segments, sums, counts
    std::vector<int> proot(static_cast<size_t>(P));
    MH_CHECK(ctx, hipMemcpyAsync(proot.data(), roots, sizeof(int) * P, hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    std::sort(proot.begin(), proot.end());
    // scratch for gathers: reuse up / up2 / comp (free after the forest)
    MH_CHECK(ctx, hipMemcpyAsync(up, proot.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    std::vector<double> hsum(3 * size_t(P));
    std::vector<unsigned> hcnt(static_cast<size_t>(P));
    std::vector<int> hlab(static_cast<size_t>(P));
    {
        DevMem mg;
        const int64_t bg = al(24 * int64_t(P)) + 2 * al(4 * int64_t(P));
        MH_CHECK(ctx, mg.alloc(bg, s));
        Arena AG{static_cast<char *>(mg.p), 0, bg};
        double *osum = AG.take<double>(3 * int64_t(P));
        unsigned *ocnt = AG.take<unsigned>(P);
        int *olab = AG.take<int>(P);
        if (!olab) return mh::set_error(ctx, MH_ERR_ARG, __func__, "gather layout");
        hipLaunchKernelGGL(gather_sums_kernel, dim3(mh::blocks_for(P)), dim3(mh::kBlock), 0, s, up, P,
                           psum, pcnt, osum, ocnt);
        // the pieces' segment labels (their roots' voxels)
        hipLaunchKernelGGL(gather_points_kernel, dim3(mh::blocks_for(P)), dim3(mh::kBlock), 0, s, up, P,
                           pw, pts, labels, dirs, (unsigned *)nullptr, olab, (float *)nullptr);
        MH_CHECK_LAUNCH(ctx, "piece gathers");
        MH_CHECK(ctx, hipMemcpyAsync(hsum.data(), osum, 24 * size_t(P), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipMemcpyAsync(hcnt.data(), ocnt, 4 * size_t(P), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipMemcpyAsync(hlab.data(), olab, 4 * size_t(P), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
    }
    for (int k = 0; k < P; ++k)
        if (hlab[size_t(k)] < 0 || hlab[size_t(k)] >= n)
            return mh::set_error(ctx, MH_ERR_ARG, __func__, "labels do not match the ridge");
    // segments: sorted distinct labels (the segments' smallest voxels)
    std::vector<int> slab(hlab);
    std::sort(slab.begin(), slab.end());
    slab.erase(std::unique(slab.begin(), slab.end()), slab.end());
    const int S = int(slab.size());
    stats[1] = S;
    std::vector<int> psg(static_cast<size_t>(P));
    for (int k = 0; k < P; ++k)
        psg[size_t(k)] = int(std::lower_bound(slab.begin(), slab.end(), hlab[size_t(k)]) - slab.begin());
    std::vector<double> ssum(3 * size_t(S), 0.0), tot(static_cast<size_t>(S), 0.0), big(static_cast<size_t>(S), 0.0);
    std::vector<int64_t> scnt(static_cast<size_t>(S), 0);
    std::vector<int> npiece(static_cast<size_t>(S), 0);
    for (int k = 0; k < P; ++k) {
        const size_t sg = size_t(psg[size_t(k)]);
        const double *v = &hsum[3 * size_t(k)];
        for (int c = 0; c < 3; ++c) ssum[3 * sg + size_t(c)] += v[c];
        const double pl = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        tot[sg] += pl;
        big[sg] = std::max(big[sg], pl);
        scnt[sg] += hcnt[size_t(k)];
        ++npiece[sg];
    }
    // the nearest-voxel transform's workspace (the largest segment's, then
    // every voxel's nearest kept point)
    DevMem m2;
    MH_CHECK(ctx, m2.alloc(al(5 * 4 * n), s));
    unsigned *edt = static_cast<unsigned *>(m2.p);
    std::vector<unsigned long long> hcuts(std::min<unsigned>(nzero, kCutCap));
    if (!hcuts.empty()) {
        MH_CHECK(ctx, hipMemcpyAsync(hcuts.data(), cuts, 8 * hcuts.size(), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
    }
    auto piece_node = [&](unsigned w) {
        const int r = int(w >> 1);
        return int(std::lower_bound(proot.begin(), proot.end(), r) - proot.begin());
    };

    // the segments' tree from the K x K tables by Boruvka rounds on the
    // device (the same unique tree as Prim's: the keys are distinct)
    auto segment_mst_device = [&](int K, const unsigned *w2min, const unsigned long long *idxmin,
                                  std::vector<std::array<int64_t, 2>> &inter) -> int {
        DevMem ms;
        const int64_t bs = 2 * al(4 * int64_t(K)) + al(4 * int64_t(K)) + 2 * al(8 * int64_t(K)) +
                           al(8 * int64_t(K)) + al(4 * int64_t(K));
        MH_CHECK(ctx, ms.alloc(bs, s));
        Arena AS{static_cast<char *>(ms.p), 0, bs};
        int *scomp = AS.take<int>(K);
        int *sup = AS.take<int>(K);
        unsigned *cw = AS.take<unsigned>(K);
        auto *cx = AS.take<unsigned long long>(K);
        auto *edge = AS.take<unsigned long long>(K);
        auto *centry = AS.take<int64_t>(K);
        int *sup2 = AS.take<int>(K);
        if (!sup2) return mh::set_error(ctx, MH_ERR_ARG, __func__, "segment tree layout");
        const unsigned nbK = mh::blocks_for(K), nbKK = mh::blocks_for(int64_t(K) * K);
        hipLaunchKernelGGL(sm_init_kernel, dim3(nbK), dim3(mh::kBlock), 0, s, K, scomp, cw, cx, edge);
        for (int round = 0;; ++round) {
            if (round >= kMaxRounds) return mh::set_error(ctx, MH_ERR_HIP, __func__, "segment tree rounds");
            hipLaunchKernelGGL(sm_w2_kernel, dim3(nbKK), dim3(mh::kBlock), 0, s, w2min, K, scomp, cw);
            hipLaunchKernelGGL(sm_pair_kernel, dim3(nbKK), dim3(mh::kBlock), 0, s, w2min, idxmin, K, scomp,
                               cw, cx);
            hipLaunchKernelGGL(sm_entry_kernel, dim3(nbKK), dim3(mh::kBlock), 0, s, w2min, idxmin, K, scomp,
                               cw, cx, centry);
            MH_CHECK(ctx, hipMemsetAsync(flags, 0, 4 * kFlagSlots * kFlagStride, s));
            hipLaunchKernelGGL(sm_hook_kernel, dim3(nbK), dim3(mh::kBlock), 0, s, K, scomp, cw, cx, centry,
                               sup, edge, flags);
            MH_CHECK_LAUNCH(ctx, "segment tree round");
            bool hooked = false;
            int rc = any_flag(hooked);
            if (rc != MH_OK) return rc;
            if (!hooked) break;
            rc = jump_until_settled(scomp, sup, sup2, K, nbK);
            if (rc != MH_OK) return rc;
            hipLaunchKernelGGL(sm_relabel_kernel, dim3(nbK), dim3(mh::kBlock), 0, s, scomp, sup, K, cw, cx);
        }
        std::vector<unsigned long long> he(static_cast<size_t>(K));
        MH_CHECK(ctx, hipMemcpyAsync(he.data(), edge, 8 * size_t(K), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
        for (unsigned long long e : he)
            if (e != kNoKey) inter.push_back({int64_t(e >> kPtBits), int64_t(e & kPtMask)});
        if (int(inter.size()) != K - 1) return mh::set_error(ctx, MH_ERR_HIP, __func__, "segment tree");
        return MH_OK;
    };

    // 5-6. The tree between the segments marked in `keep` and the walk over
    // it: psign[k] receives the sign of piece k's root (0 for pieces of
    // dropped segments), so that a point's sign is psign of its piece times
    // its parity. Sets kept_pts and seg_of_root (the kept segments' numbers).
    std::vector<char> keep(static_cast<size_t>(S), 1);
    std::vector<float> psign(static_cast<size_t>(P), 0.f);
    int64_t kept_pts = 0;
    auto tree_walk = [&]() -> int {
        std::vector<int> kseg;
        for (int k = 0; k < S; ++k)
            if (keep[size_t(k)]) kseg.push_back(k);
        const int K = int(kseg.size());
        int bigk = 0;
        kept_pts = 0;
        for (int t = 0; t < K; ++t) {
            kept_pts += scnt[size_t(kseg[size_t(t)])];
            if (scnt[size_t(kseg[size_t(t)])] > scnt[size_t(kseg[size_t(bigk)])]) bigk = t;
        }
        stats[2] = K;
        stats[3] = kept_pts;
        if (K == 0 || kept_pts <= kDensePoints || K > kMaxKept) return MH_DEFER;
        const int64_t m_others = kept_pts - scnt[size_t(kseg[size_t(bigk)])];
        if (m_others > kMaxOthers) return MH_DEFER;

        // each kept segment's smallest point (the point of its label voxel,
        // a ridge voxel), marked with the segment's number
        std::vector<int> seg_root_pt(static_cast<size_t>(K));
        {
            std::vector<int> labs(static_cast<size_t>(K)), ids(static_cast<size_t>(K));
            for (int t = 0; t < K; ++t) labs[size_t(t)] = slab[size_t(kseg[size_t(t)])];
            std::iota(ids.begin(), ids.end(), 0);
            DevMem mkk;
            MH_CHECK(ctx, mkk.alloc(3 * al(4 * int64_t(K)), s));
            Arena AKK{static_cast<char *>(mkk.p), 0, 3 * al(4 * int64_t(K))};
            int *d_labs = AKK.take<int>(K), *d_pts = AKK.take<int>(K), *d_ids = AKK.take<int>(K);
            if (!d_ids) return mh::set_error(ctx, MH_ERR_ARG, __func__, "segment root layout");
            MH_CHECK(ctx, hipMemcpyAsync(d_labs, labs.data(), 4 * size_t(K), hipMemcpyHostToDevice, s));
            MH_CHECK(ctx, hipMemcpyAsync(d_ids, ids.data(), 4 * size_t(K), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(gather_int_kernel, dim3(mh::blocks_for(K)), dim3(mh::kBlock), 0, s, d_labs, K, at,
                               d_pts);
            hipLaunchKernelGGL(fill_kernel<int>, dim3(nbR), dim3(mh::kBlock), 0, s, seg_of_root, int64_t(R), -1);
            hipLaunchKernelGGL(scatter_int_kernel, dim3(mh::blocks_for(K)), dim3(mh::kBlock), 0, s, d_pts, d_ids,
                               K, seg_of_root);
            MH_CHECK_LAUNCH(ctx, "kept segment roots");
            MH_CHECK(ctx, hipMemcpyAsync(seg_root_pt.data(), d_pts, 4 * size_t(K), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipStreamSynchronize(s));
            for (int t = 0; t < K; ++t)
                if (seg_root_pt[size_t(t)] < 0)
                    return mh::set_error(ctx, MH_ERR_ARG, __func__, "segment label off the ridge");
        }

        // the lightest point pair of every two segments, then the segments'
        // minimum spanning tree (Prim over the K x K table; the keys are
        // distinct, so it is the unique tree under their order)
        std::vector<std::array<int64_t, 2>> inter;  // (a, b) point indices
        if (K > 1) {
            DevMem mt;
            const int64_t bt = 2 * al(4 * int64_t(K + 1)) + al(4 * m_others) + al(4 * int64_t(K) * K) +
                               al(8 * int64_t(K) * K);
            MH_CHECK(ctx, mt.alloc(bt, s));
            Arena AT{static_cast<char *>(mt.p), 0, bt};
            unsigned *start = AT.take<unsigned>(K + 1);
            unsigned *fill = AT.take<unsigned>(K + 1);
            int *others = AT.take<int>(m_others);
            unsigned *w2min = AT.take<unsigned>(int64_t(K) * K);
            auto *idxmin = AT.take<unsigned long long>(int64_t(K) * K);
            if (!idxmin) return mh::set_error(ctx, MH_ERR_ARG, __func__, "tree layout");
            std::vector<unsigned> hstart(static_cast<size_t>(K) + 1, 0);
            for (int t = 0; t < K; ++t)
                hstart[size_t(t) + 1] =
                    hstart[size_t(t)] + (t == bigk ? 0u : unsigned(scnt[size_t(kseg[size_t(t)])]));
            MH_CHECK(ctx, hipMemcpyAsync(start, hstart.data(), 4 * (size_t(K) + 1), hipMemcpyHostToDevice, s));
            MH_CHECK(ctx, hipMemsetAsync(fill, 0, 4 * (size_t(K) + 1), s));
            hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, s, edt, n,
                               kNone);
            hipLaunchKernelGGL(split_kept_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, pts, labels, at,
                               seg_of_root, R, bigk, edt, fill, start, others, cnt + 4);
            MH_CHECK_LAUNCH(ctx, "split_kept_kernel");
            unsigned bad = 0;
            MH_CHECK(ctx, hipMemcpyAsync(&bad, cnt + 4, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipStreamSynchronize(s));
            if (bad) return mh::set_error(ctx, MH_ERR_HIP, __func__, "segment buckets overflow");
            const int rc = mh_nearest_marked(ctx, dims, edt, s);
            if (rc != MH_OK) return rc;
            hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(mh::blocks_for(int64_t(K) * K)), dim3(mh::kBlock), 0,
                               s, w2min, int64_t(K) * K, kNone);
            hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(mh::blocks_for(int64_t(K) * K)),
                               dim3(mh::kBlock), 0, s, idxmin, int64_t(K) * K, (unsigned long long)kNoKey);
            const int m = int(m_others);
            const unsigned nbp = mh::blocks_for(int64_t(m) * K);
            const unsigned *ed = edt + n, *eb = edt + 2 * n;
            hipLaunchKernelGGL(pair_kernel<0>, dim3(nbp), dim3(mh::kBlock), 0, s, others, m, pts, labels, at,
                               seg_of_root, start, K, bigk, ed, eb, g, w2min, idxmin);
            hipLaunchKernelGGL(pair_kernel<1>, dim3(nbp), dim3(mh::kBlock), 0, s, others, m, pts, labels, at,
                               seg_of_root, start, K, bigk, ed, eb, g, w2min, idxmin);
            MH_CHECK_LAUNCH(ctx, "segment pairs");
            if (K > kHostPrimMax) {
                const int rc2 = segment_mst_device(K, w2min, idxmin, inter);
                if (rc2 != MH_OK) return rc2;
            } else {
            std::vector<unsigned> hw2(static_cast<size_t>(K) * K);
            std::vector<unsigned long long> hix(static_cast<size_t>(K) * K);
            MH_CHECK(ctx, hipMemcpyAsync(hw2.data(), w2min, 4 * hw2.size(), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipMemcpyAsync(hix.data(), idxmin, 8 * hix.size(), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipStreamSynchronize(s));
            auto key_less = [](unsigned w0, unsigned long long x0, unsigned w1, unsigned long long x1) {
                return w0 != w1 ? w0 < w1 : x0 < x1;
            };
            auto slot = [K](int a, int b) { return size_t(std::min(a, b)) * size_t(K) + size_t(std::max(a, b)); };
            std::vector<char> in(static_cast<size_t>(K), 0);
            std::vector<unsigned> bw(static_cast<size_t>(K), kNone);
            std::vector<unsigned long long> bx(static_cast<size_t>(K), kNoKey);
            in[0] = 1;
            for (int j = 1; j < K; ++j) {
                bw[size_t(j)] = hw2[slot(0, j)];
                bx[size_t(j)] = hix[slot(0, j)];
            }
            for (int step = 1; step < K; ++step) {
                int j = -1;
                for (int k = 0; k < K; ++k)
                    if (!in[size_t(k)] && (j < 0 || key_less(bw[size_t(k)], bx[size_t(k)], bw[size_t(j)], bx[size_t(j)])))
                        j = k;
                if (j < 0 || bx[size_t(j)] == kNoKey)
                    return mh::set_error(ctx, MH_ERR_HIP, __func__, "segment pair without a distance");
                in[size_t(j)] = 1;
                inter.push_back({int64_t(bx[size_t(j)] >> kPtBits), int64_t(bx[size_t(j)] & kPtMask)});
                for (int k = 0; k < K; ++k) {
                    if (in[size_t(k)]) continue;
                    const size_t sl = slot(j, k);
                    if (key_less(hw2[sl], hix[sl], bw[size_t(k)], bx[size_t(k)])) {
                        bw[size_t(k)] = hw2[sl];
                        bx[size_t(k)] = hix[sl];
                    }
                }
            }
            }
        }

        // the walk over pieces. Endpoints of the tree's inter-segment edges
        // and of the kept segments' zero-product forest edges, with their pieces.
        const int root_pt = *std::min_element(seg_root_pt.begin(), seg_root_pt.end());
        std::vector<int> qlist;
        for (const auto &e : inter) {
            qlist.push_back(int(e[0]));
            qlist.push_back(int(e[1]));
        }
        for (unsigned long long c : hcuts) {
            qlist.push_back(int((c >> kPtBits) & kPtMask));
            qlist.push_back(int(c & kPtMask));
        }
        qlist.push_back(root_pt);
        const int Q = int(qlist.size());
        std::vector<unsigned> qpw(static_cast<size_t>(Q));
        std::vector<int> qlab(static_cast<size_t>(Q));
        std::vector<float> qdir(3 * size_t(Q));
        {
            DevMem mq;
            const int64_t bq = 3 * al(4 * int64_t(Q)) + al(12 * int64_t(Q));
            MH_CHECK(ctx, mq.alloc(bq, s));
            Arena AQ{static_cast<char *>(mq.p), 0, bq};
            int *dl = AQ.take<int>(Q);
            unsigned *opw = AQ.take<unsigned>(Q);
            int *olab = AQ.take<int>(Q);
            float *odir = AQ.take<float>(3 * int64_t(Q));
            if (!odir) return mh::set_error(ctx, MH_ERR_ARG, __func__, "gather layout");
            MH_CHECK(ctx, hipMemcpyAsync(dl, qlist.data(), 4 * size_t(Q), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(gather_points_kernel, dim3(mh::blocks_for(Q)), dim3(mh::kBlock), 0, s, dl, Q, pw,
                               pts, labels, dirs, opw, olab, odir);
            MH_CHECK_LAUNCH(ctx, "gather_points_kernel");
            MH_CHECK(ctx, hipMemcpyAsync(qpw.data(), opw, 4 * size_t(Q), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipMemcpyAsync(qlab.data(), olab, 4 * size_t(Q), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipMemcpyAsync(qdir.data(), odir, 12 * size_t(Q), hipMemcpyDeviceToHost, s));
            MH_CHECK(ctx, hipStreamSynchronize(s));
        }
        auto is_kept_label = [&](int lab) {
            const int sg = int(std::lower_bound(slab.begin(), slab.end(), lab) - slab.begin());
            return sg < S && slab[size_t(sg)] == lab && keep[size_t(sg)];
        };
        // final pieces: forest pieces joined across the tree's inter-segment
        // edges with a non-zero product; zero products become cuts (the
        // child end keeps its direction)
        HostSets fp(P);
        struct WalkCut {
            unsigned pwa, pwb;
        };
        std::vector<WalkCut> wcuts;
        for (size_t e = 0; e < inter.size(); ++e) {
            const size_t qa = 2 * e, qb = qa + 1;
            const float d = dot3(&qdir[3 * qa], &qdir[3 * qb]);
            if (d != d) return MH_DEFER;
            if (d == 0.f) {
                wcuts.push_back(WalkCut{qpw[qa], qpw[qb]});
                continue;
            }
            const unsigned char rel = d < 0.f ? 1 : 0;
            fp.join(piece_node(qpw[qa]), piece_node(qpw[qb]),
                    (unsigned char)((qpw[qa] & 1u) ^ (qpw[qb] & 1u) ^ rel));
        }
        for (size_t c = 0; c < hcuts.size(); ++c) {
            const size_t qa = 2 * inter.size() + 2 * c, qb = qa + 1;
            if (!is_kept_label(qlab[qa])) continue;  // a cut in a dropped segment
            wcuts.push_back(WalkCut{qpw[qa], qpw[qb]});
        }
        // breadth-first over the final pieces through the cuts, from the root's
        std::vector<int> gstart(static_cast<size_t>(P) + 1, 0);
        std::vector<int> ga(wcuts.size()), gb(wcuts.size());
        for (size_t c = 0; c < wcuts.size(); ++c) {
            unsigned char t;
            ga[c] = fp.find(piece_node(wcuts[c].pwa), t);
            gb[c] = fp.find(piece_node(wcuts[c].pwb), t);
            ++gstart[size_t(ga[c]) + 1];
            ++gstart[size_t(gb[c]) + 1];
        }
        for (int k = 0; k < P; ++k) gstart[size_t(k) + 1] += gstart[size_t(k)];
        struct Adj {
            int group;
            int cut;
            bool b_side;  // entering through the cut's b end
        };
        std::vector<Adj> adj(2 * wcuts.size());
        {
            std::vector<int> f(gstart.begin(), gstart.end() - 1);
            for (size_t c = 0; c < wcuts.size(); ++c) {
                adj[size_t(f[size_t(ga[c])]++)] = Adj{gb[c], int(c), true};
                adj[size_t(f[size_t(gb[c])]++)] = Adj{ga[c], int(c), false};
            }
        }
        std::vector<int> anchor_q(static_cast<size_t>(P), -2);  // -1: the root point; else cut * 2 (+1: b end)
        std::vector<int> queue;
        const unsigned root_pw = qpw[size_t(Q - 1)];
        unsigned char tpar;
        const int g0 = fp.find(piece_node(root_pw), tpar);
        anchor_q[size_t(g0)] = -1;
        queue.push_back(g0);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int gcur = queue[h];
            for (int t = gstart[size_t(gcur)]; t < gstart[size_t(gcur) + 1]; ++t) {
                const Adj &e = adj[size_t(t)];
                if (anchor_q[size_t(e.group)] != -2) continue;
                anchor_q[size_t(e.group)] = 2 * e.cut + (e.b_side ? 1 : 0);
                queue.push_back(e.group);
            }
        }
        // sign of every kept piece's root: (-1)^(par(piece root -> group
        // root) ^ par(anchor -> group root)); the anchor keeps its direction
        std::fill(psign.begin(), psign.end(), 0.f);
        for (int k = 0; k < P; ++k) {
            if (!keep[size_t(psg[size_t(k)])]) continue;
            unsigned char pk;
            const int gk = fp.find(k, pk);
            const int aq = anchor_q[size_t(gk)];
            if (aq == -2) return mh::set_error(ctx, MH_ERR_HIP, __func__, "piece not reached by the walk");
            const unsigned apw =
                aq == -1 ? root_pw : ((aq & 1) ? wcuts[size_t(aq >> 1)].pwb : wcuts[size_t(aq >> 1)].pwa);
            unsigned char pa;
            (void)fp.find(piece_node(apw), pa);
            const unsigned par_anchor = (apw & 1u) ^ pa;  // anchor -> its piece root -> group root
            psign[size_t(k)] = ((pk ^ par_anchor) & 1u) ? -1.f : 1.f;
        }
        return MH_OK;
    };

    // The kept segments. First by bounds: each segment's length lies in
    // [lo, hi] over its pieces' open relative signs, widened by a margin for
    // the summation order; decided when the bounds separate the kept from
    // the rest. Otherwise by the signs of the full tree's walk (as the host
    // construction does), deciding unless two lengths at the boundary agree
    // within rounding.
    const int want = std::min(S, int(std::ceil(3.0 * std::log(M_E + double(S)))));
    if (clean && want < S) {
        std::vector<double> lo(static_cast<size_t>(S)), hi(static_cast<size_t>(S));
        for (int k = 0; k < S; ++k) {
            const double *v = &ssum[3 * size_t(k)];
            const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const double margin = 1e-9 * (tot[size_t(k)] + 1.0);
            if (npiece[size_t(k)] > 1) {
                lo[size_t(k)] = std::max(0.0, 2.0 * big[size_t(k)] - tot[size_t(k)]) - margin;
                hi[size_t(k)] = tot[size_t(k)] + margin;
            } else {
                lo[size_t(k)] = len - margin;
                hi[size_t(k)] = len + margin;
            }
        }
        std::vector<double> shi(hi), slo(lo);
        std::sort(shi.begin(), shi.end());
        std::sort(slo.begin(), slo.end());
        bool decided = true;
        int nin = 0;
        for (int k = 0; k < S && decided; ++k) {
            // others whose upper bound reaches this one's lower bound; others
            // whose lower bound exceeds this one's upper bound
            const int64_t reach = (shi.end() - std::lower_bound(shi.begin(), shi.end(), lo[size_t(k)])) - 1;
            const int64_t beat =
                int64_t(slo.size()) - (std::upper_bound(slo.begin(), slo.end(), hi[size_t(k)]) - slo.begin());
            const bool sure_in = reach < want, sure_out = beat >= want;
            decided = sure_in != sure_out;
            keep[size_t(k)] = sure_in ? 1 : 0;
            nin += sure_in ? 1 : 0;
        }
        if (!decided || nin != want) {
            const auto t_full = std::chrono::steady_clock::now();
            std::fill(keep.begin(), keep.end(), 1);
            const int rc = tree_walk();
            stats[5] = usec_since(t_full);
            if (rc != MH_OK) return rc;
            std::vector<double> len(static_cast<size_t>(S), 0.0), acc(3 * size_t(S), 0.0);
            for (int k = 0; k < P; ++k)
                for (int c = 0; c < 3; ++c)
                    acc[3 * size_t(psg[size_t(k)]) + size_t(c)] += double(psign[size_t(k)]) * hsum[3 * size_t(k) + size_t(c)];
            for (int k = 0; k < S; ++k) {
                const double *v = &acc[3 * size_t(k)];
                len[size_t(k)] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
            std::vector<int> order(static_cast<size_t>(S));
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return len[size_t(x)] > len[size_t(y)]; });
            const double a = len[size_t(order[size_t(want - 1)])], b = len[size_t(order[size_t(want)])];
            if (a - b <= 1e-9 * (a + 1.0)) return MH_DEFER;  // a tie within rounding at the boundary
            std::fill(keep.begin(), keep.end(), 0);
            for (int t = 0; t < want; ++t) keep[size_t(order[size_t(t)])] = 1;
        }
    }
    {
        const int rc = tree_walk();
        if (rc != MH_OK) return rc;
    }
    std::vector<int> kroots;
    std::vector<float> ksign;
    for (int k = 0; k < P; ++k)
        if (keep[size_t(psg[size_t(k)])]) {
            kroots.push_back(proot[size_t(k)]);
            ksign.push_back(psign[size_t(k)]);
        }

    // 7. signed kept points, the kept ridge, every voxel's sign
    const int NK = int(kroots.size());
    {
        DevMem mk;
        const int64_t bk = al(4 * int64_t(NK)) + al(4 * int64_t(NK)) + al(4 * int64_t(ntile)) + al(4 * 8) +
                           al(4 * kept_pts) + al(6 * kept_pts) + al(12 * kept_pts);
        MH_CHECK(ctx, mk.alloc(bk, s));
        Arena AK{static_cast<char *>(mk.p), 0, bk};
        int *dr = AK.take<int>(NK);
        float *ds = AK.take<float>(NK);
        const unsigned ptile = unsigned((int64_t(R) + kScanTile - 1) / kScanTile);
        unsigned *pc = AK.take<unsigned>(ptile);
        unsigned *ptot = AK.take<unsigned>(8);
        int *kept = AK.take<int>(kept_pts);
        auto *kc = AK.take<uint16_t>(3 * kept_pts);
        float *kd = AK.take<float>(3 * kept_pts);
        if (!kd) return mh::set_error(ctx, MH_ERR_ARG, __func__, "output layout");
        MH_CHECK(ctx, hipMemcpyAsync(dr, kroots.data(), 4 * size_t(NK), hipMemcpyHostToDevice, s));
        MH_CHECK(ctx, hipMemcpyAsync(ds, ksign.data(), 4 * size_t(NK), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scatter_sign_kernel, dim3(mh::blocks_for(NK)), dim3(mh::kBlock), 0, s, dr, ds, NK,
                           fsign);
        hipLaunchKernelGGL(fill_kernel<float>, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, s, kept_out, n, 0.f);
        // the count first: the compaction writes one entry per flagged point
        hipLaunchKernelGGL(kept_flags_kernel, dim3(nbR), dim3(mh::kBlock), 0, s, pts, labels, at, seg_of_root,
                           R, kflag, kept_out);
        hipLaunchKernelGGL(flag_count_kernel<unsigned char>, dim3(ptile), dim3(mh::kBlock), 0, s, kflag,
                           int64_t(R), pc);
        hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, pc, ptile, ptot);
        MH_CHECK_LAUNCH(ctx, "kept points");
        unsigned nkept = 0;
        MH_CHECK(ctx, hipMemcpyAsync(&nkept, ptot, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        MH_CHECK(ctx, hipStreamSynchronize(s));
        if (int64_t(nkept) != kept_pts) return mh::set_error(ctx, MH_ERR_HIP, __func__, "kept point count");
        hipLaunchKernelGGL(compact_kernel<unsigned char>, dim3(ptile), dim3(mh::kBlock), 0, s, kflag,
                           int64_t(R), pc, kept, (int *)nullptr);
        hipLaunchKernelGGL(kept_out_kernel, dim3(mh::blocks_for(kept_pts)), dim3(mh::kBlock), 0, s, kept,
                           int(kept_pts), pts, pw, fsign, dirs, g, kc, kd);
        MH_CHECK_LAUNCH(ctx, "kept ridge");
        const float scaling[3] = {float(voxel_size[0]), float(voxel_size[1]), float(voxel_size[2])};
        const int rc = mh_nonridge_signs(ctx, dirs, dims, kc, kd, kept_pts, scaling, edt,
                                         5 * 4 * n, stream);
        if (rc != MH_OK) return rc;
        MH_CHECK(ctx, hipStreamSynchronize(s));
    }
    stats[6] = usec_since(t_start);
    return MH_OK;
}
