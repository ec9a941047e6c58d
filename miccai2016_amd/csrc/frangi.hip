// Frangi vesselness on gfx950: separable Gaussian-derivative passes, the
// per-voxel 3x3 Jacobi eigensolver and the multi-scale vesselness update.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mh_common.h"
#include "mh_reduce.h"

namespace {

// ------------------------------------------------------ separable passes ---

constexpr int kMaxRadius = 63;
constexpr int kMaxTaps = 2 * kMaxRadius + 1;
constexpr int kTapsPad = 128;   // taps per order (>= kMaxTaps), zero beyond K

struct ConvOut {
    float *ptr;
    int order;   // which tap set
    int stride;  // elements between consecutive voxels (1 = plain field, 6 = AoS x6)
    int offset;  // element offset of this component
};

struct ConvGroup {
    const float *in;
    int nout;
    ConvOut out[3];
};

// One pass along one axis; each group reads one input field and feeds up to
// three outputs (derivative orders), so intermediates shared by several
// Hessian entries are computed once.
struct ConvPass {
    ConvGroup grp[6];
    int ngroups;
    int radius;
    // taps of each derivative order in convolution order: ctaps[o][j] is tap
    // K-1-j, the weight of element p - R + j for output p; zero for j >= K
    float ctaps[3][kTapsPad];
};

// Index of position p of an axis of length len, mirror-completed and tiled
// with period 2 (len - 1) (scipy's 'mirror' mode; a length-1 axis repeats its
// one element).
__device__ __forceinline__ int mirror_index(int p, int len)
{
    if (p >= 0 && p < len) return p;
    const int period = 2 * (len - 1);
    if (period == 0) return 0;
    if (p < 0 && -p < len) return -p;                 // one reflection at 0
    if (p >= len && p < period) return period - p;    // one reflection at len - 1
    int t = p % period;
    t += t < 0 ? period : 0;
    return t < len ? t : period - t;
}

// The axis-2 pass stages a tile of the input line segments plus a halo of R
// taps on each side in LDS (each element loaded once, coalesced along axis
// 2); the axis-0/1 passes read their columns straight from global memory.
// Every output sums its K taps in the order j = 0 .. K-1 of the true
// convolution (tap K-1-j against element p - R + j). The taps are uniform
// across the work-group and are staged from the kernel arguments into LDS.

constexpr int kTileX = 64;    // consecutive voxels along axis 2 per tile
constexpr int kRunA = 8;      // outputs per lane along the pass axis (axes 0, 1)
constexpr int kColWaves = 4;  // waves (runs) per work-group of the axis-0/1 pass
constexpr int kRows2 = 16;    // rows (axis 1) per tile of the axis-2 pass
// LDS row width of the axis-2 pass: the tile's 64 + 2R inputs and 8 zero
// columns, rounded to 16 bytes; each of the row's 16 threads loads every
// 16th entry. Rows of up to 16 kRowLoads floats (R <= 26) prefetch.
constexpr int kRowLoads = 8;

// Pass along axis 0 or 1. A wave owns 64 consecutive voxels along axis 2
// (one per lane) and a run of kRunA (8) consecutive outputs along the pass
// axis; a work-group holds kColWaves runs, one above the other, and reads
// one input group. Each lane slides a window of the 8 column elements its
// outputs need for the current tap along its column, as a ring indexed by
// position mod 8: at tap j, output q reads element q + j from slot
// (q + j) mod 8, and once all 8 outputs have used element j its slot takes
// element j + 8, loaded from global memory. Taps go in unrolled blocks of 8
// (the last one cut at K by a uniform test), so every slot index is a
// constant. The column's rows are a uniform address plus the lane's voxel,
// so a load costs one address add; the runs' overlapping reads (K + 7 rows
// for 8 outputs) hit in the caches, and HBM sees each element about once.
// There is no barrier after the taps are staged: each wave's loads, taps and
// stores go on by themselves, which keeps more loads in flight than a tile
// staged in LDS, whose load, compute and store phases alternate (144 against
// 184 us for the axis-1 pass at 256^3, bit for bit the same outputs).
template <int NOUT, bool INTERIOR>
__device__ __forceinline__ void col_run(const float *wts, const float *in, const ConvOut *out,
                                        int R, int K, bool live, int64_t cbase, int xs,
                                        int64_t sa, int a0, int len)
{
    // element m of the column is in[cbase + m * sa + xs]; cbase is uniform,
    // xs the lane's voxel (a valid one for lanes past the volume, which
    // store nothing); INTERIOR: every element of the run's window, a0 - R
    // .. a0 + R + 8, lies inside the axis
    const int lo = a0 - R;
    auto fetch = [&](int e) -> float {
        const int m = INTERIOR ? lo + e : mirror_index(lo + e, len);
        return in[cbase + int64_t(m) * sa + xs];
    };
    float acc[NOUT][kRunA];
#pragma unroll
    for (int q = 0; q < kRunA; ++q)
#pragma unroll
        for (int k = 0; k < NOUT; ++k) acc[k][q] = 0.f;
    float win[kRunA];
#pragma unroll
    for (int q = 0; q < kRunA; ++q) win[q] = fetch(q);
    auto tap = [&](int j, int jj) {
        float w[NOUT];
#pragma unroll
        for (int k = 0; k < NOUT; ++k) w[k] = wts[k * kTapsPad + j];
#pragma unroll
        for (int q = 0; q < kRunA; ++q) {
            const float v = win[(jj + q) & (kRunA - 1)];
#pragma unroll
            for (int k = 0; k < NOUT; ++k) acc[k][q] += w[k] * v;
        }
        // element j is used up: its slot takes element j + kRunA
        win[jj] = fetch(j + kRunA);
    };
    int j0 = 0;
    for (; j0 + kRunA <= K; j0 += kRunA) {
#pragma unroll
        for (int jj = 0; jj < kRunA; ++jj) tap(j0 + jj, jj);
    }
    const int rem = K - j0;
#pragma unroll
    for (int jj = 0; jj < kRunA; ++jj)
        if (jj < rem) tap(j0 + jj, jj);
    if (!live) return;
#pragma unroll
    for (int q = 0; q < kRunA; ++q) {
        const int a = a0 + q;
        if (a < len) {
            const int64_t i = cbase + int64_t(a) * sa + xs;
#pragma unroll
            for (int k = 0; k < NOUT; ++k) out[k].ptr[i * out[k].stride + out[k].offset] = acc[k][q];
        }
    }
}

template <int AXIS>
__global__ void __launch_bounds__(64 * kColWaves)
conv_col_kernel(const ConvPass P, int d0, int d1, int d2)
{
    const int len = AXIS == 0 ? d0 : d1;
    const int other = AXIS == 0 ? d1 : d0;
    const int n2 = (d2 + kTileX - 1) / kTileX;
    const int na = (len + kRunA * kColWaves - 1) / (kRunA * kColWaves);
    // one input group per work-group, the groups of a run side by side
    unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int g = int(b % unsigned(P.ngroups));
    b /= unsigned(P.ngroups);
    const int t2 = int(b % unsigned(n2));
    const int ta = int((b / unsigned(n2)) % unsigned(na));
    const int o = int(b / (unsigned(n2) * unsigned(na)));
    if (o >= other) return;   // uniform per work-group
    const int x = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const ConvGroup &G = P.grp[g];
    __shared__ float wts[3 * kTapsPad];   // the group's taps, [output][j]
    for (int t = threadIdx.x; t < G.nout * kTapsPad; t += 64 * kColWaves)
        wts[t] = P.ctaps[G.out[t / kTapsPad].order][t % kTapsPad];
    __syncthreads();
    const int a0 = (ta * kColWaves + wv) * kRunA;
    if (a0 >= len) return;    // uniform per wave; no barrier follows
    const int R = P.radius, K = 2 * R + 1;
    const int64_t sa = AXIS == 0 ? int64_t(d1) * d2 : int64_t(d2);
    const int64_t cbase = (AXIS == 0 ? int64_t(o) * d2 : int64_t(o) * d1 * d2) + int64_t(t2) * kTileX;
    const bool live = t2 * kTileX + x < d2;
    const int xs = live ? x : d2 - 1 - t2 * kTileX;
    const bool interior = a0 - R >= 0 && a0 + R + kRunA < len;
    const float *in = G.in;
    const ConvOut *out = G.out;
    if (interior) {
        if (G.nout == 3) col_run<3, true>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
        else if (G.nout == 2) col_run<2, true>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
        else col_run<1, true>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
    } else {
        if (G.nout == 3) col_run<3, false>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
        else if (G.nout == 2) col_run<2, false>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
        else col_run<1, false>(wts, in, out, R, K, live, cbase, xs, sa, a0, len);
    }
}

// Pass along axis 2 (contiguous), one output per group. Tile: 64 outputs
// along axis 2 x 16 rows along axis 1; the tile rows are loaded by 16
// threads each. Lane l of wave w computes the 4 consecutive outputs
// 4 (l mod 16) .. 4 (l mod 16) + 3 of row 4 w + l / 16. LDS: 16 rows of
// W floats, W = 64 + 2R rounded up past 8 zero columns to a multiple of 4
// (16-byte aligned rows). The thread's window is two groups of 4 elements,
// A (elements j0 .. j0+3 of its run) and B (j0+4 .. j0+7), each one 16-byte
// LDS read: taps go in blocks of 4 (the last one cut at K by a uniform
// test), where output q at tap j0 + jj reads element q + jj of A:B, then A
// takes B and B the next 4 elements. The results of all groups stay in
// registers and are stored together at the end.
__device__ __forceinline__ float4 lds4(const float *p)
{
    return *reinterpret_cast<const float4 *>(p);
}

__device__ __forceinline__ void row_tap(float *a, const float4 &A, const float4 &B, int jj, float w)
{
    const float v[8] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] += w * v[q + jj];
}

template <int NL>
__global__ void __launch_bounds__(256)
conv_row_kernel(const ConvPass P, int d0, int d1, int d2)
{
    extern __shared__ float lds[];
    const int n2 = (d2 + kTileX - 1) / kTileX;
    const int n1 = (d1 + kRows2 - 1) / kRows2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int t2 = int(b % unsigned(n2));
    const int t1 = int((b / unsigned(n2)) % unsigned(n1));
    const int i0 = int(b / (unsigned(n2) * unsigned(n1)));
    if (i0 >= d0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = wave * 4 + (lane >> 4), xq = 4 * (lane & 15);
    const int R = P.radius, K = 2 * R + 1, valid = kTileX + 2 * R;
    const int width = (valid + 8 + 3) & ~3;
    const int s0 = t2 * kTileX;
    // loading: row lr of the tile, columns lc, lc + 16, ...
    const int lr = threadIdx.x / 16, lc = threadIdx.x % 16;
    const int li1 = t1 * kRows2 + lr;
    const int64_t lrow = (int64_t(i0) * d1 + li1) * d2;
    const float *line = lds + row * width + xq;
    // NL > 0 (rows of at most 16 NL floats): the next group's row segment is
    // fetched into registers while the current group's taps run, so the
    // loads' latency hides behind them
    float nxt[NL > 0 ? NL : 1];
    auto fetch = [&](int g) {
        const float *in = P.grp[g].in;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int c = lc + 16 * k;
            nxt[k] = (li1 < d1 && c < valid) ? in[lrow + mirror_index(s0 - R + c, d2)] : 0.f;
        }
    };
    if (NL > 0) fetch(0);
    float res[4][6];
#pragma unroll
    for (int g = 0; g < 6; ++g) {
        if (g >= P.ngroups) break;
        const ConvGroup &G = P.grp[g];
        if (NL > 0) {
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int c = lc + 16 * k;
                if (c < width) lds[lr * width + c] = nxt[k];
            }
        } else {
            for (int c = lc; c < width; c += 16)
                lds[lr * width + c] =
                    (li1 < d1 && c < valid) ? G.in[lrow + mirror_index(s0 - R + c, d2)] : 0.f;
        }
        __syncthreads();
        if (NL > 0 && g + 1 < P.ngroups) fetch(g + 1);
        float a[4] = {0.f, 0.f, 0.f, 0.f};
        const float *wts = P.ctaps[G.out[0].order];
        float4 A = lds4(line), B = lds4(line + 4);
        int j0 = 0;
#pragma unroll 2
        for (; j0 + 4 <= K; j0 += 4) {
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) row_tap(a, A, B, jj, wts[j0 + jj]);
            A = B;
            B = lds4(line + j0 + 8);
        }
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
            if (j0 + jj < K) row_tap(a, A, B, jj, wts[j0 + jj]);
#pragma unroll
        for (int q = 0; q < 4; ++q) res[q][g] = a[q];
        __syncthreads();
    }
    // the Hessian's AoS x6 output (all six groups into one 16-byte aligned
    // buffer, component g at offset g): the tile's block goes through LDS
    // (free after the last group's barrier) and leaves row by row in 16-byte
    // pieces in lane order, so a wave's store covers 1 KiB of consecutive
    // output; a thread storing its own 4 voxels' 24 floats scatters each
    // store over 6 KiB (222 against 199 us at 256^3)
    bool aos6 = P.ngroups == 6;
#pragma unroll
    for (int g = 0; g < 6; ++g)
        aos6 = aos6 && P.grp[g].out[0].ptr == P.grp[0].out[0].ptr &&
               P.grp[g].out[0].stride == 6 && P.grp[g].out[0].offset == g;
    aos6 = aos6 && (reinterpret_cast<uintptr_t>(P.grp[0].out[0].ptr) & 15) == 0;
    if (aos6) {   // uniform per work-group
        constexpr int kRowFloats = 6 * kTileX, kRowPieces = kRowFloats / 4;
        float4 *sp = reinterpret_cast<float4 *>(lds + row * kRowFloats + xq * 6);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = res[(4 * k + e) / 6][(4 * k + e) % 6];
            sp[k] = make_float4(v[0], v[1], v[2], v[3]);
        }
        __syncthreads();
        const int nf = 6 * min(kTileX, d2 - s0);   // valid floats per row
        float *ob = P.grp[0].out[0].ptr;
        for (int f = threadIdx.x; f < kRows2 * kRowPieces; f += 256) {
            const int r = f / kRowPieces, e0 = 4 * (f % kRowPieces);
            const int ri = t1 * kRows2 + r;
            if (ri >= d1 || e0 >= nf) continue;
            const float4 v = *reinterpret_cast<const float4 *>(lds + r * kRowFloats + e0);
            float *dst = ob + ((int64_t(i0) * d1 + ri) * d2 + s0) * 6 + e0;
            if (e0 + 4 <= nf) {
                *reinterpret_cast<float4 *>(dst) = v;
            } else {
                const float vv[4] = {v.x, v.y, v.z, v.w};
                for (int e = 0; e < 4 && e0 + e < nf; ++e) dst[e] = vv[e];
            }
        }
        return;
    }
    const int i1 = t1 * kRows2 + row;
    if (i1 >= d1) return;
    const int64_t i = (int64_t(i0) * d1 + i1) * d2 + s0 + xq;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (s0 + xq + q >= d2) continue;
#pragma unroll
        for (int g = 0; g < 6; ++g) {
            if (g >= P.ngroups) break;
            const ConvOut &O = P.grp[g].out[0];
            O.ptr[(i + q) * O.stride + O.offset] = res[q][g];
        }
    }
}

int launch_pass(mh_ctx *ctx, int axis, const ConvPass &P, const int64_t *dims, hipStream_t s)
{
    // no output may alias an input (tiles read halos that other tiles would
    // overwrite); the axis-2 pass keeps one output per group in registers
    MH_REQUIRE(ctx, P.ngroups >= 1 && P.ngroups <= 6, "1 to 6 groups per pass");
    for (int g = 0; g < P.ngroups; ++g) {
        MH_REQUIRE(ctx, P.grp[g].nout >= 1 && P.grp[g].nout <= (axis == 2 ? 1 : 3),
                   "outputs per group: 1 to 3 (1 along axis 2)");
        for (int h = 0; h < P.ngroups; ++h)
            for (int u = 0; u < P.grp[h].nout; ++u)
                MH_REQUIRE(ctx, P.grp[h].out[u].ptr != P.grp[g].in, "a pass cannot work in place");
    }
    const int d0 = int(dims[0]), d1 = int(dims[1]), d2 = int(dims[2]);
    const int64_t n2 = (d2 + kTileX - 1) / kTileX;
    if (axis == 2) {
        const int64_t blocks = n2 * ((d1 + kRows2 - 1) / kRows2) * d0;
        MH_REQUIRE(ctx, blocks < (int64_t(1) << 31), "volume too large");
        // the input tile, or the AoS x6 block of the results
        const size_t shm = sizeof(float) * kRows2 * std::max((kTileX + 2 * P.radius + 8 + 3) & ~3, 6 * kTileX);
        const int width = (kTileX + 2 * P.radius + 8 + 3) & ~3;
        if (width <= 16 * kRowLoads)
            hipLaunchKernelGGL(conv_row_kernel<kRowLoads>, dim3(unsigned(blocks)), dim3(256), shm, s, P, d0,
                               d1, d2);
        else
            hipLaunchKernelGGL(conv_row_kernel<0>, dim3(unsigned(blocks)), dim3(256), shm, s, P, d0, d1, d2);
    } else {
        const int len = axis == 0 ? d0 : d1, other = axis == 0 ? d1 : d0;
        const int64_t blocks = n2 * ((len + kRunA * kColWaves - 1) / (kRunA * kColWaves)) * other *
                               P.ngroups;
        MH_REQUIRE(ctx, blocks < (int64_t(1) << 31), "volume too large");
        if (axis == 0)
            hipLaunchKernelGGL(conv_col_kernel<0>, dim3(unsigned(blocks)), dim3(64 * kColWaves), 0, s,
                               P, d0, d1, d2);
        else
            hipLaunchKernelGGL(conv_col_kernel<1>, dim3(unsigned(blocks)), dim3(64 * kColWaves), 0, s,
                               P, d0, d1, d2);
    }
    MH_CHECK_LAUNCH(ctx, "conv pass");
    return MH_OK;
}

void set_taps(ConvPass &P, int slot, const float *taps, int r)
{
    // taps of a shorter arm are centred in the common radius P.radius, then
    // stored in convolution order (tap K-1-j at j), zero beyond K
    float centred[kMaxTaps] = {};
    const int shift = P.radius - r, K = 2 * P.radius + 1;
    for (int k = 0; k < 2 * r + 1; ++k) centred[k + shift] = taps[k];
    for (int j = 0; j < kTapsPad; ++j) P.ctaps[slot][j] = j < K ? centred[K - 1 - j] : 0.f;
}

// ------------------------------------------------------------ eigensolver ---

// Rotations that re-order the eigenvector rows into ascending |lambda|
// order, scalar last, indexed by (order[0] << 1) | (order[1] > order[2]).
// Each is the signed permutation matrix with row k = e_order[k], negated as a
// whole for odd permutations (det = +1), converted to a unit quaternion.
#define MH_SQ1_2 0.70710678118654752440f
__constant__ float kSorter[6][4] = {
    {0.f, 0.f, 0.f, 1.f},
    {0.f, -MH_SQ1_2, MH_SQ1_2, 0.f},
    {-MH_SQ1_2, MH_SQ1_2, 0.f, 0.f},
    {0.5f, 0.5f, 0.5f, -0.5f},
    {0.5f, 0.5f, 0.5f, 0.5f},
    {-MH_SQ1_2, 0.f, MH_SQ1_2, 0.f},
};

__device__ __forceinline__ void atomic_max_nonneg(float *addr, float v)
{
    // non-negative floats order like their bit patterns
    atomicMax(reinterpret_cast<unsigned *>(addr), __float_as_uint(v));
}

struct MaxNonNeg {
    __device__ static float identity() { return 0.f; }
    __device__ static float combine(float a, float b) { return fmaxf(a, b); }
};

// Cyclic Jacobi on one symmetric 3x3 matrix with the approximate Givens
// rotation of McAdams et al. (UW-Madison TR 1690, 2011, algorithm 2):
// gamma = 3 + 2 sqrt 2 decides between the exact half-angle pair and the
// fixed (cos, sin)(pi/8). Convergence (sum |offdiag| <= max|input| * eps) is
// tested before every sweep, so a lane that has converged stops rotating.
__global__ void __launch_bounds__(mh::kBlock)
eig_kernel(const float *in6, float *out6, int64_t n, int maxiter, float *norm_max)
{
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    float nrm = 0.f;
    if (i < n) {
        const float *h = in6 + 6 * i;
        float d[3] = {h[0], h[1], h[2]};
        float o[3] = {h[5], h[4], h[3]};  // o[r]: element (p, q), r = 3 - p - q
        float qv[3] = {0.f, 0.f, 0.f};
        float qw = 1.f;
        float tol = 0.f;
        for (int k = 0; k < 3; ++k) tol = fmaxf(tol, fmaxf(fabsf(d[k]), fabsf(o[k])));
        tol *= FLT_EPSILON;
        const float kGamma = 5.82842712474619009760f;
        const float kCos = 0.92387953251128675613f, kSin = 0.38268343236508977173f;
        bool conv = false;
        for (int it = 0; it < maxiter; ++it) {
            if (fabsf(o[0]) + fabsf(o[1]) + fabsf(o[2]) <= tol) {
                conv = true;
                break;
            }
#pragma unroll
            for (int pq = 0; pq < 3; ++pq) {
                const int p = pq == 2 ? 1 : 0;
                const int q = pq == 0 ? 1 : 2;
                const int r = 3 - p - q;
                float c = 2.f * (d[q] - d[p]);
                float s = o[r];
                const float s2 = s * s, c2 = c * c;
                const bool exact = kGamma * s2 < c2;
                const float w = rsqrtf(s2 + c2);
                c = exact ? w * c : kCos;
                s = exact ? w * s : kSin;
                // accumulate the half-angle rotation into the quaternion
                const float ap = c * qv[p] - s * qv[q];
                const float aq = c * qv[q] + s * qv[p];
                qv[p] = ap;
                qv[q] = aq;
                const float sg = r == 1 ? s : -s;
                const float nw = c * qw + sg * qv[r];
                const float nr = c * qv[r] - sg * qw;
                qw = nw;
                qv[r] = nr;
                // full-angle rotation J M J^T of the matrix
                const float cc = 1.f - 2.f * (s * s);
                const float ss = 2.f * c * s;
                const float t1 = o[r] * cc + d[p] * ss;
                const float t2 = d[q] * cc + o[r] * ss;
                d[p] = cc * (d[p] * cc - o[r] * ss) - ss * (o[r] * cc - d[q] * ss);
                d[q] = cc * t2 + ss * t1;
                o[r] = cc * t1 - ss * t2;
                const float oq = o[q];
                o[q] = cc * oq - ss * o[p];
                o[p] = cc * o[p] + ss * oq;
            }
        }
        float res[6];
        if (conv) {
            // compare-exchange (0,1), (1,2), (0,1) on |lambda|, tracking order
            int ord0 = 0, ord1 = 1, ord2 = 2;
            float l0 = d[0], l1 = d[1], l2 = d[2];
            if (!(fabsf(l0) < fabsf(l1))) { float t = l0; l0 = l1; l1 = t; int u = ord0; ord0 = ord1; ord1 = u; }
            if (!(fabsf(l1) < fabsf(l2))) { float t = l1; l1 = l2; l2 = t; int u = ord1; ord1 = ord2; ord2 = u; }
            if (!(fabsf(l0) < fabsf(l1))) { float t = l0; l0 = l1; l1 = t; int u = ord0; ord0 = ord1; ord1 = u; }
            const float *L = kSorter[(ord0 << 1) | (ord1 > ord2 ? 1 : 0)];
            // Hamilton product L * q (scalar last), then transpose (negate scalar)
            const float sw = L[3] * qw - (L[0] * qv[0] + L[1] * qv[1] + L[2] * qv[2]);
            const float sx = L[3] * qv[0] + L[0] * qw + L[1] * qv[2] - L[2] * qv[1];
            const float sy = L[3] * qv[1] - L[0] * qv[2] + L[1] * qw + L[2] * qv[0];
            const float sz = L[3] * qv[2] + L[0] * qv[1] - L[1] * qv[0] + L[2] * qw;
            float sq[3];
            mh::squeeze_quat(-sw, sx, sy, sz, sq);
            res[0] = l0; res[1] = l1; res[2] = l2;
            res[3] = sq[0]; res[4] = sq[1]; res[5] = sq[2];
            nrm = sqrtf(l0 * l0 + l1 * l1 + l2 * l2);
            if (!(nrm >= 0.f)) nrm = 0.f;  // NaN is neutral in the maximum
        } else {
            for (int k = 0; k < 6; ++k) res[k] = __builtin_nanf("");
        }
        float *e = out6 + 6 * i;
        for (int k = 0; k < 6; ++k) e[k] = res[k];
    }
    if (norm_max) {
        nrm = mh::block_reduce<MaxNonNeg>(nrm);
        if (threadIdx.x == 0 && nrm > 0.f) atomic_max_nonneg(norm_max, nrm);
    }
}

__global__ void __launch_bounds__(mh::kBlock)
eig_norm_kernel(const float *__restrict__ e6, int64_t n, float *norm_max)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float nrm = 0.f;
    if (i < n) {
        const float *e = e6 + 6 * i;
        nrm = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        if (!(nrm >= 0.f)) nrm = 0.f;
    }
    nrm = mh::block_reduce<MaxNonNeg>(nrm);
    if (threadIdx.x == 0 && nrm > 0.f) atomic_max_nonneg(norm_max, nrm);
}

// ------------------------------------------------------------- vesselness ---

__global__ void __launch_bounds__(mh::kBlock)
vesselness_kernel(float *__restrict__ val, float *__restrict__ quat, const float *__restrict__ e6,
                  int64_t n, float alpha, float beta, float c_value, const float *norm_max,
                  int bright)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float c = c_value >= 0.f ? c_value : 0.5f * (*norm_max);
    const float *e = e6 + 6 * i;
    const float l1 = fabsf(e[0]);
    float l2 = bright ? e[1] : -e[1];
    float l3 = bright ? e[2] : -e[2];
    const bool valid = l2 < 0.f && l3 < 0.f;
    l2 = fabsf(l2);
    l3 = fabsf(l3);
    const float rB = l1 / (sqrtf(l2 * l3) + FLT_MIN);
    const float rA = l2 / (l3 + FLT_MIN);
    const float S2 = l1 * l1 + l2 * l2 + l3 * l3;
    const float ra = rA / alpha, rb = rB / beta;
    const float v = float(valid) * ((1.f - expf(-0.5f * (ra * ra))) * expf(-0.5f * (rb * rb)) *
                                    (1.f - expf(-0.5f * S2 / (c * c))));
    if (v > val[i]) {
        val[i] = v;
        quat[3 * i] = e[3];
        quat[3 * i + 1] = e[4];
        quat[3 * i + 2] = e[5];
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI ---

MH_EXPORT int64_t mh_hessian_work_floats(const int64_t dims[3])
{
    if (!mh::dims_ok(dims)) return -1;
    return 9 * dims[0] * dims[1] * dims[2];
}

MH_EXPORT int mh_gauss_filter(mh_ctx *ctx, const float *in, float *out, float *swap,
                              const int64_t dims[3], const float *taps0, int r0,
                              const float *taps1, int r1, const float *taps2, int r2, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, in && out && swap && taps0 && taps1 && taps2, "null pointer");
    const int r[3] = {r0, r1, r2};
    const float *taps[3] = {taps0, taps1, taps2};
    for (int a = 0; a < 3; ++a) MH_REQUIRE(ctx, r[a] >= 0 && r[a] <= kMaxRadius, "radius out of range");
    // axis 0 -> swap? choose buffers so the result of the third pass lands in out:
    // pass0: in -> out, pass1: out -> swap, pass2: swap -> out
    const float *src[3] = {in, out, swap};
    float *dst[3] = {out, swap, out};
    for (int a = 0; a < 3; ++a) {
        ConvPass P;
        std::memset(&P, 0, sizeof(P));
        P.radius = r[a];
        set_taps(P, 0, taps[a], r[a]);
        P.ngroups = 1;
        P.grp[0].in = src[a];
        P.grp[0].nout = 1;
        P.grp[0].out[0] = ConvOut{dst[a], 0, 1, 0};
        int rc = launch_pass(ctx, a, P, dims, (hipStream_t)stream);
        if (rc) return rc;
    }
    return MH_OK;
}

MH_EXPORT int mh_gauss_hessian(mh_ctx *ctx, const float *img, const int64_t dims[3],
                               const float *taps, const int radius[3], float *hess6, float *work,
                               void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, img && taps && radius && hess6 && work, "null pointer");
    int rmax = 0;
    for (int a = 0; a < 3; ++a) {
        MH_REQUIRE(ctx, radius[a] >= 0 && radius[a] <= kMaxRadius, "radius out of range");
        rmax = std::max(rmax, radius[a]);
    }
    const int64_t n = dims[0] * dims[1] * dims[2];
    const int K = 2 * rmax + 1;
    auto tapset = [&](int axis, int order) { return taps + (3 * axis + order) * K; };
    float *A[3] = {work, work + n, work + 2 * n};  // axis-0 results for orders 0, 1, 2
    float *B = work + 3 * n;                          // 6 axis-1 results
    hipStream_t s = (hipStream_t)stream;

    // pass along axis 0: image -> A0, A1, A2
    {
        ConvPass P;
        std::memset(&P, 0, sizeof(P));
        P.radius = radius[0];
        for (int o = 0; o < 3; ++o) set_taps(P, o, tapset(0, o), radius[0]);
        P.ngroups = 1;
        P.grp[0] = ConvGroup{img, 3, {{A[0], 0, 1, 0}, {A[1], 1, 1, 0}, {A[2], 2, 1, 0}}};
        int rc = launch_pass(ctx, 0, P, dims, s);
        if (rc) return rc;
    }
    // pass along axis 1: B0..B5 = A0*o0, A0*o1, A0*o2, A1*o0, A1*o1, A2*o0
    {
        ConvPass P;
        std::memset(&P, 0, sizeof(P));
        P.radius = radius[1];
        for (int o = 0; o < 3; ++o) set_taps(P, o, tapset(1, o), radius[1]);
        P.ngroups = 3;
        P.grp[0] = ConvGroup{A[0], 3, {{B, 0, 1, 0}, {B + n, 1, 1, 0}, {B + 2 * n, 2, 1, 0}}};
        P.grp[1] = ConvGroup{A[1], 2, {{B + 3 * n, 0, 1, 0}, {B + 4 * n, 1, 1, 0}, {}}};
        P.grp[2] = ConvGroup{A[2], 1, {{B + 5 * n, 0, 1, 0}, {}, {}}};
        int rc = launch_pass(ctx, 1, P, dims, s);
        if (rc) return rc;
    }
    // pass along axis 2 into AoS x6 [H00, H11, H22, H01, H02, H12]:
    //   H00 = (o2,o0,o0) from B5, H11 = (o0,o2,o0) from B2, H22 = (o0,o0,o2) from B0,
    //   H01 = (o1,o1,o0) from B4, H02 = (o1,o0,o1) from B3, H12 = (o0,o1,o1) from B1
    {
        ConvPass P;
        std::memset(&P, 0, sizeof(P));
        P.radius = radius[2];
        for (int o = 0; o < 3; ++o) set_taps(P, o, tapset(2, o), radius[2]);
        P.ngroups = 6;
        P.grp[0] = ConvGroup{B + 5 * n, 1, {{hess6, 0, 6, 0}, {}, {}}};
        P.grp[1] = ConvGroup{B + 2 * n, 1, {{hess6, 0, 6, 1}, {}, {}}};
        P.grp[2] = ConvGroup{B, 1, {{hess6, 2, 6, 2}, {}, {}}};
        P.grp[3] = ConvGroup{B + 4 * n, 1, {{hess6, 0, 6, 3}, {}, {}}};
        P.grp[4] = ConvGroup{B + 3 * n, 1, {{hess6, 1, 6, 4}, {}, {}}};
        P.grp[5] = ConvGroup{B + n, 1, {{hess6, 1, 6, 5}, {}, {}}};
        int rc = launch_pass(ctx, 2, P, dims, s);
        if (rc) return rc;
    }
    return MH_OK;
}

MH_EXPORT int mh_eigs(mh_ctx *ctx, const float *hess6, float *eig6, int64_t n, int maxiter,
                      float *norm_max, void *stream)
{
    MH_REQUIRE(ctx, hess6 && eig6 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    MH_REQUIRE(ctx, maxiter >= 0, "negative maxiter");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(eig_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, hess6, eig6, n, maxiter, norm_max);
    MH_CHECK_LAUNCH(ctx, "eig_kernel");
    return MH_OK;
}

MH_EXPORT int mh_update_vesselness(mh_ctx *ctx, float *val, float *quat, const float *eig6,
                                   int64_t n, float alpha, float beta, float c_value,
                                   const float *norm_max, int bright, void *stream)
{
    MH_REQUIRE(ctx, val && quat && eig6 && n >= 0, "bad arguments");
    MH_REQUIRE(ctx, c_value >= 0.f || norm_max, "c < 0 needs norm_max_dev");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(vesselness_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, val, quat, eig6, n, alpha, beta, c_value, norm_max,
                       bright);
    MH_CHECK_LAUNCH(ctx, "vesselness_kernel");
    return MH_OK;
}

MH_EXPORT int mh_eig_norm_max(mh_ctx *ctx, const float *eig6, int64_t n, float *norm_max,
                              void *stream)
{
    MH_REQUIRE(ctx, eig6 && norm_max && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    int rc = mh_fill(ctx, norm_max, 1, 0.f, stream);
    if (rc || n == 0) return rc;
    hipLaunchKernelGGL(eig_norm_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, eig6, n, norm_max);
    MH_CHECK_LAUNCH(ctx, "eig_norm_kernel");
    return MH_OK;
}
