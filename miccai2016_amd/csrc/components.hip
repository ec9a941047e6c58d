// Face-connected components of a thresholded labelling on gfx950, and the
// Dice of the largest one (the experiment's dice_for, phantom_experiment.py:
// 50-103, which labels with scipy's default structure).
//
// Union-find over voxel indices: every foreground voxel starts as its own
// root; each voxel joins its backward face neighbours (axis 0, 1, 2) by
// hooking the larger root under the smaller one with an atomic minimum, so
// parent indices only ever decrease and every find and retry ends. After a
// final compression every voxel holds the smallest index of its component:
// the labels do not depend on the order the hooks happened in, and ordering
// components by that index is scipy's label order (raster order of their
// first voxel), so ties of the largest size go to the first component, as
// the reference's strict comparison does.
#include "mh_common.h"

namespace {

constexpr int kBg = -1;

__device__ __forceinline__ int find_root(const int *__restrict__ lab, int x)
{
    int p = lab[x];
    while (p != x) {
        x = p;
        p = lab[x];
    }
    return x;
}

// Join the trees of a and b. Whenever an atomic minimum finds a's parent
// already changed (to `old`), the loop goes on to join old with b, so no
// link that a hook overwrote is lost.
__device__ __forceinline__ void unite(int *__restrict__ lab, int a, int b)
{
    while (true) {
        a = find_root(lab, a);
        b = find_root(lab, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        // hook root a (the larger) under b
        const int old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}

__global__ void __launch_bounds__(mh::kBlock)
cc_init_kernel(const float *__restrict__ u, float thr, int64_t n, int *__restrict__ lab,
               int *__restrict__ cnt)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lab[i] = u[i] >= thr ? int(i) : kBg;
    cnt[i] = 0;
}

__global__ void __launch_bounds__(mh::kBlock)
cc_merge_kernel(int *__restrict__ lab, int d0, int d1, int d2)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] == kBg) return;
    const int i2 = int(i % d2), i1 = int((i / d2) % d1), i0 = int(i / (int64_t(d2) * d1));
    const int64_t s0 = int64_t(d1) * d2;
    if (i2 > 0 && lab[i - 1] != kBg) unite(lab, int(i), int(i - 1));
    if (i1 > 0 && lab[i - d2] != kBg) unite(lab, int(i), int(i - d2));
    if (i0 > 0 && lab[i - s0] != kBg) unite(lab, int(i), int(i - s0));
}

__global__ void __launch_bounds__(mh::kBlock)
cc_compress_kernel(int *__restrict__ lab, int64_t n, int *__restrict__ cnt)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] == kBg) return;
    const int r = find_root(lab, int(i));
    atomicAdd(&cnt[r], 1);
    lab[i] = r;  // parents stay ancestors, so concurrent finds remain valid
}

// Largest component: per block the maximum of (size << 32 | ~root) over the
// roots (a larger size wins, then the smaller root), then one block.
__global__ void __launch_bounds__(mh::kBlock)
cc_largest_partials_kernel(const int *__restrict__ cnt, int64_t n,
                           unsigned long long *__restrict__ part)
{
    unsigned long long best = 0;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int c = cnt[i];
        if (c > 0) {
            const unsigned long long key =
                (static_cast<unsigned long long>(unsigned(c)) << 32) | unsigned(~unsigned(i));
            best = key > best ? key : best;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    __shared__ unsigned long long red[mh::kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < mh::kBlock / 64; ++w) best = red[w] > best ? red[w] : best;
        part[blockIdx.x] = best;
    }
}

// out[0] = |L| (or -1 without foreground), out[1] = |L & T|, out[2] = |T|,
// out[3] = the root of L; part[0..nparts) holds the largest-key partials.
__global__ void __launch_bounds__(mh::kBlock)
cc_dice_kernel(const int *__restrict__ lab, const uint8_t *__restrict__ truth, int64_t n,
               const unsigned long long *__restrict__ part, int nparts,
               unsigned long long *__restrict__ counts_part)
{
    __shared__ unsigned long long key_s;
    if (threadIdx.x == 0) {
        unsigned long long k = 0;
        for (int q = 0; q < nparts; ++q) k = part[q] > k ? part[q] : k;
        key_s = k;
    }
    __syncthreads();
    const unsigned long long key = key_s;
    const int root = key ? int(~unsigned(key & 0xffffffffull)) : -2;
    unsigned inter = 0, t = 0;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const bool tr = truth[i] != 0;
        t += tr;
        inter += tr && lab[i] == root;
    }
    for (int off = 32; off > 0; off >>= 1) {
        inter += __shfl_xor(inter, off);
        t += __shfl_xor(t, off);
    }
    __shared__ unsigned red[2][mh::kBlock / 64];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = inter;
        red[1][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned a = 0, b = 0;
        for (int w = 0; w < mh::kBlock / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        counts_part[2 * blockIdx.x] = a;
        counts_part[2 * blockIdx.x + 1] = b;
        if (blockIdx.x == 0) counts_part[2 * gridDim.x] = key;
    }
}

__global__ void cc_dice_final_kernel(const unsigned long long *__restrict__ counts_part, int nparts,
                                     long long *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long a = 0, b = 0;
    for (int q = 0; q < nparts; ++q) {
        a += counts_part[2 * q];
        b += counts_part[2 * q + 1];
    }
    const unsigned long long key = counts_part[2 * nparts];
    out[0] = key ? static_cast<long long>(key >> 32) : -1;
    out[1] = static_cast<long long>(a);
    out[2] = static_cast<long long>(b);
    out[3] = key ? static_cast<long long>(~unsigned(key & 0xffffffffull)) : -1;
}

constexpr unsigned kParts = 1024;

int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

// ----------------------------------------------- ridge segments (26-conn.) --
// The ridge's hysteresis and its segments for the cleaning step, as union-find
// over the 26-neighbourhood: every candidate voxel joins its 13 backward
// neighbours. A component is kept when it holds a seed (a candidate whose
// response reaches the upper threshold), which is the flood fill from the
// seeds through the candidates of ridge_detection.py:80-131.

__global__ void __launch_bounds__(mh::kBlock)
ridge_init_kernel(const float *__restrict__ cand, int64_t n, int *__restrict__ lab,
                  int *__restrict__ seeded)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lab[i] = cand[i] != 0.f ? int(i) : kBg;
    seeded[i] = 0;
}

__global__ void __launch_bounds__(mh::kBlock)
ridge_merge26_kernel(int *__restrict__ lab, int d0, int d1, int d2)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] == kBg) return;
    const int i2 = int(i % d2), i1 = int((i / d2) % d1), i0 = int(i / (int64_t(d2) * d1));
    // backward neighbours: (-1, *, *), (0, -1, *), (0, 0, -1)
    for (int a = -1; a <= 0; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c) {
                if (a == 0 && (b > 0 || (b == 0 && c >= 0))) continue;
                const int j0 = i0 + a, j1 = i1 + b, j2 = i2 + c;
                if (j0 < 0 || j1 < 0 || j1 >= d1 || j2 < 0 || j2 >= d2) continue;
                const int64_t j = (int64_t(j0) * d1 + j1) * d2 + j2;
                if (lab[j] != kBg) unite(lab, int(i), int(j));
            }
}

// final labels (smallest voxel index of the component); a component holding
// a candidate whose response is >= upper is marked at its root
__global__ void __launch_bounds__(mh::kBlock)
ridge_compress_kernel(int *__restrict__ lab, const float *__restrict__ resp, float upper,
                      int use_upper, int64_t n, int *__restrict__ seeded)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] == kBg) return;
    const int r = find_root(lab, int(i));
    lab[i] = r;  // parents stay ancestors, so concurrent finds remain valid
    if (!use_upper || resp[i] >= upper) seeded[r] = 1;
}

__global__ void __launch_bounds__(mh::kBlock)
ridge_keep_kernel(int *__restrict__ lab, const int *__restrict__ seeded, int64_t n,
                  float *__restrict__ ridge)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = lab[i];
    const bool keep = r != kBg && seeded[r] != 0;
    ridge[i] = keep ? 1.f : 0.f;
    lab[i] = keep ? r : kBg;
}

}  // namespace

MH_EXPORT int64_t mh_largest_component_work_bytes(const int64_t dims[3])
{
    if (!mh::dims_ok(dims)) return -1;
    const int64_t n = dims[0] * dims[1] * dims[2];
    return 2 * align256(sizeof(int) * n) + align256(sizeof(unsigned long long) * kParts) +
           align256(sizeof(unsigned long long) * (2 * kParts + 1));
}

MH_EXPORT int mh_largest_component_dice(mh_ctx *ctx, const float *u, float threshold,
                                        const uint8_t *truth, const int64_t dims[3],
                                        int *labels_dev, void *work, int64_t work_bytes,
                                        long long *out_dev, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims) && u && truth && work && out_dev, "bad arguments");
    MH_REQUIRE(ctx, work_bytes >= mh_largest_component_work_bytes(dims), "workspace too small");
    MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(work) & 255) == 0, "workspace must be 256-byte aligned");
    const int64_t n = dims[0] * dims[1] * dims[2];
    char *w = static_cast<char *>(work);
    int *lab = labels_dev ? labels_dev : reinterpret_cast<int *>(w);
    int *cnt = reinterpret_cast<int *>(w + align256(sizeof(int) * n));
    auto *part = reinterpret_cast<unsigned long long *>(w + 2 * align256(sizeof(int) * n));
    auto *cpart = reinterpret_cast<unsigned long long *>(
        w + 2 * align256(sizeof(int) * n) + align256(sizeof(unsigned long long) * kParts));
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = mh::blocks_for(n);
    const unsigned np = std::min<unsigned>(kParts, nb);
    hipLaunchKernelGGL(cc_init_kernel, dim3(nb), dim3(mh::kBlock), 0, s, u, threshold, n, lab, cnt);
    hipLaunchKernelGGL(cc_merge_kernel, dim3(nb), dim3(mh::kBlock), 0, s, lab, int(dims[0]),
                       int(dims[1]), int(dims[2]));
    hipLaunchKernelGGL(cc_compress_kernel, dim3(nb), dim3(mh::kBlock), 0, s, lab, n, cnt);
    hipLaunchKernelGGL(cc_largest_partials_kernel, dim3(np), dim3(mh::kBlock), 0, s, cnt, n, part);
    hipLaunchKernelGGL(cc_dice_kernel, dim3(np), dim3(mh::kBlock), 0, s, lab, truth, n, part, int(np),
                       cpart);
    hipLaunchKernelGGL(cc_dice_final_kernel, dim3(1), dim3(64), 0, s, cpart, int(np), out_dev);
    MH_CHECK_LAUNCH(ctx, "largest component");
    return MH_OK;
}

MH_EXPORT int64_t mh_ridge_segments_work_bytes(const int64_t dims[3])
{
    if (!mh::dims_ok(dims)) return -1;
    return align256(sizeof(int) * dims[0] * dims[1] * dims[2]);
}

MH_EXPORT int mh_ridge_segments(mh_ctx *ctx, const float *cand, const float *resp, float upper,
                                int use_upper, const int64_t dims[3], float *ridge, int *labels,
                                void *work, int64_t work_bytes, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, cand && ridge && labels && work && (resp || !use_upper), "null pointer");
    const int64_t n = dims[0] * dims[1] * dims[2];
    MH_REQUIRE(ctx, work_bytes >= mh_ridge_segments_work_bytes(dims), "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    int *seeded = static_cast<int *>(work);
    const unsigned nb = mh::blocks_for(n);
    hipLaunchKernelGGL(ridge_init_kernel, dim3(nb), dim3(mh::kBlock), 0, s, cand, n, labels, seeded);
    hipLaunchKernelGGL(ridge_merge26_kernel, dim3(nb), dim3(mh::kBlock), 0, s, labels, int(dims[0]),
                       int(dims[1]), int(dims[2]));
    hipLaunchKernelGGL(ridge_compress_kernel, dim3(nb), dim3(mh::kBlock), 0, s, labels, resp, upper,
                       use_upper, n, seeded);
    hipLaunchKernelGGL(ridge_keep_kernel, dim3(nb), dim3(mh::kBlock), 0, s, labels, seeded, n, ridge);
    MH_CHECK_LAUNCH(ctx, "ridge segment kernels");
    return MH_OK;
}
