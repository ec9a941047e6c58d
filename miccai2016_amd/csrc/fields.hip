// Per-voxel fields around the max-flow on gfx950: terminal capacities,
// |grad I|^2, the nonterminal weight, min-max normalisation, response-weighted
// directions, and a synthetic helix phantom generator for load tests.
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

// cs, ct from the log-likelihood ratio of N(1, s^2) vs N(0, s^2) with a
// regulariser eps, scaled by 1 / log((p_hat + eps) / eps) (the reference's
// normpdf-based weights, computed here in FP32).
__global__ void __launch_bounds__(mh::kBlock)
capacities_kernel(const float *__restrict__ img, int64_t n, float inv_m2var, float norm,
                  float eps, float rq, float *__restrict__ cs, float *__restrict__ ct)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = img[i];
    const float bgp = expf(x * x * inv_m2var) * norm;
    const float fgp = expf((x - 1.f) * (x - 1.f) * inv_m2var) * norm;
    const float r = logf((fgp + eps) / (bgp + eps));
    cs[i] = rq * fmaxf(r, 0.f);
    ct[i] = rq * fmaxf(-r, 0.f);
}

// |grad I|^2 with central differences inside and one-sided differences at
// the faces (numpy.gradient, edge order 1).
__global__ void __launch_bounds__(mh::kBlock)
grad_mag_sq_kernel(const float *__restrict__ img, int d0, int d1, int d2, float *__restrict__ out)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int i2 = int(i % d2), i1 = int((i / d2) % d1), i0 = int(i / (int64_t(d2) * d1));
    const int c[3] = {i0, i1, i2};
    const int dim[3] = {d0, d1, d2};
    const int64_t st[3] = {int64_t(d1) * d2, d2, 1};
    float acc = 0.f;
    for (int a = 0; a < 3; ++a) {
        float g;
        if (dim[a] < 2) g = 0.f;
        else if (c[a] == 0) g = img[i + st[a]] - img[i];
        else if (c[a] == dim[a] - 1) g = img[i] - img[i - st[a]];
        else g = (img[i + st[a]] - img[i - st[a]]) * 0.5f;
        acc += g * g;
    }
    out[i] = acc;
}

__global__ void __launch_bounds__(mh::kBlock)
nonterminal_kernel(const float *__restrict__ gms, int64_t n, float w, float k, float *__restrict__ alpha)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    alpha[i] = w * expf(-gms[i] * k);
}

// Maximum that propagates NaN (numpy.max semantics).
struct NanMaxOp {
    __device__ static float identity() { return -INFINITY; }
    __device__ static float combine(float a, float b) { return (a != a || b != b) ? NAN : fmaxf(a, b); }
};

template <class Op>
__global__ void __launch_bounds__(mh::kBlock)
reduce_partials_kernel(const float *__restrict__ x, int64_t n, float *__restrict__ part)
{
    float acc = Op::identity();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        acc = Op::combine(acc, x[i]);
    acc = mh::block_reduce<Op>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <class Op>
__global__ void __launch_bounds__(mh::kBlock) reduce_final_kernel(float *__restrict__ part, int nparts)
{
    float acc = Op::identity();
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) acc = Op::combine(acc, part[k]);
    acc = mh::block_reduce<Op>(acc);
    if (threadIdx.x == 0) part[nparts] = acc;
}

// u0 = img / (max(img) + FLT_MIN), the division correctly rounded in FP32.
__global__ void __launch_bounds__(mh::kBlock)
initial_u_kernel(const float *__restrict__ img, int64_t n, const float *__restrict__ mx,
                 float *__restrict__ u)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u[i] = img[i] / (mx[0] + FLT_MIN);
}

// Dice counts of S labellings (u >= 0.5) against one mask: per block
// |S|, |S & T| and |T| as exact integers.
__global__ void __launch_bounds__(mh::kBlock)
dice_partials_kernel(const float *__restrict__ u, int64_t n, const uint8_t *__restrict__ truth,
                     unsigned *__restrict__ part)
{
    const int s = blockIdx.y;
    const float *us = u + int64_t(s) * n;
    unsigned a = 0, both = 0, t = 0;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const bool in = us[i] >= 0.5f;
        const bool tr = truth[i] != 0;
        a += in;
        both += in && tr;
        t += tr;
    }
    __shared__ unsigned red[3][mh::kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        both += __shfl_xor(both, off);
        t += __shfl_xor(t, off);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = a;
        red[1][wid] = both;
        red[2][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned acc = 0;
        for (int w = 0; w < mh::kBlock / 64; ++w) acc += red[threadIdx.x][w];
        part[(int64_t(s) * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = acc;
    }
}

__global__ void __launch_bounds__(64)
dice_final_kernel(const unsigned *__restrict__ part, int nparts, unsigned long long *__restrict__ out)
{
    const int s = blockIdx.x;
    unsigned long long acc[3] = {0, 0, 0};
    for (int k = threadIdx.x; k < nparts; k += 64)
        for (int c = 0; c < 3; ++c) acc[c] += part[(int64_t(s) * nparts + k) * 3 + c];
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; ++c) acc[c] += __shfl_xor(acc[c], off);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) out[3 * s + c] = acc[c];
}

struct MinOp {
    __device__ static float identity() { return INFINITY; }
    __device__ static float combine(float a, float b) { return fminf(a, b); }
};
struct MaxOp {
    __device__ static float identity() { return -INFINITY; }
    __device__ static float combine(float a, float b) { return fmaxf(a, b); }
};

__global__ void __launch_bounds__(mh::kBlock)
minmax_partials_kernel(const float *__restrict__ x, int64_t n, float *__restrict__ part)
{
    float lo = INFINITY, hi = -INFINITY;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        lo = fminf(lo, x[i]);
        hi = fmaxf(hi, x[i]);
    }
    lo = mh::block_reduce<MinOp>(lo);
    hi = mh::block_reduce<MaxOp>(hi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

__global__ void __launch_bounds__(mh::kBlock)
minmax_final_kernel(float *__restrict__ part, int nparts)
{
    float lo = INFINITY, hi = -INFINITY;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        lo = fminf(lo, part[2 * k]);
        hi = fmaxf(hi, part[2 * k + 1]);
    }
    lo = mh::block_reduce<MinOp>(lo);
    hi = mh::block_reduce<MaxOp>(hi);
    if (threadIdx.x == 0) {
        part[2 * nparts] = lo;
        part[2 * nparts + 1] = hi;
    }
}

// out = (x - min) / (max - min), or 0 everywhere if max == min.
__global__ void __launch_bounds__(mh::kBlock)
normalize_kernel(const float *__restrict__ x, int64_t n, const float *__restrict__ mm,
                 float *__restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float lo = mm[0], range = mm[1] - mm[0];
    out[i] = range == 0.f ? 0.f : (1.f / range) * (x[i] - lo);
}

// v (AoS x3) scaled by the response: the GVF input field (SoA x3).
__global__ void __launch_bounds__(mh::kBlock)
weighted_dirs_kernel(const float *__restrict__ dirs, const float *__restrict__ resp, int64_t n,
                     float *__restrict__ v3)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = resp[i];
    v3[i] = dirs[3 * i] * r;
    v3[n + i] = dirs[3 * i + 1] * r;
    v3[2 * n + i] = dirs[3 * i + 2] * r;
}

// SoA x3 -> AoS x3.
__global__ void __launch_bounds__(mh::kBlock)
soa_to_aos_kernel(const float *__restrict__ soa, int64_t n, float *__restrict__ aos)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] = soa[i];
    aos[3 * i + 1] = soa[n + i];
    aos[3 * i + 2] = soa[2 * n + i];
}

// ---- synthetic helix phantom (load generation) ---------------------------
// Tube of radius r_tube around x = cx + Rh cos(2 pi z / P), y = cy + Rh sin(..),
// z along axis 0; noise N(0, sigma^2) from a counter-based hash of (seed, i)
// (SplitMix64) and the Box-Muller transform.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void __launch_bounds__(mh::kBlock)
helix_kernel(int d0, int d1, int d2, float r_tube, float r_helix, float pitch, float sigma,
             uint64_t seed, float *__restrict__ img, uint8_t *__restrict__ truth,
             float *__restrict__ tangent)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = float(i / (int64_t(d2) * d1));
    const float y = float((i / d2) % d1), x = float(i % d2);
    const float cy = 0.5f * (d1 - 1), cx = 0.5f * (d2 - 1);
    const float w = 6.28318530717958647692f / pitch;
    // closest point on the centre line: best of a sampling of t in
    // [z - r_tube - 2, z + r_tube + 2], refined by Newton steps
    const float reach = r_tube + 2.f;
    float bt = z, bd = INFINITY;
    for (float t = z - reach; t <= z + reach; t += 0.25f) {
        float s, c;
        sincosf(w * t, &s, &c);
        const float ex = cx + r_helix * c - x, ey = cy + r_helix * s - y, ez = t - z;
        const float d = ex * ex + ey * ey + ez * ez;
        if (d < bd) { bd = d; bt = t; }
    }
    for (int k = 0; k < 3; ++k) {
        float s, c;
        sincosf(w * bt, &s, &c);
        const float rx = x - (cx + r_helix * c), ry = y - (cy + r_helix * s), rz = z - bt;
        const float dx = -r_helix * w * s, dy = r_helix * w * c;
        const float g = -(rx * dx + ry * dy + rz);
        const float h = dx * dx + dy * dy + 1.f + rx * r_helix * w * w * c + ry * r_helix * w * w * s;
        if (fabsf(h) > 1e-6f) bt -= g / h;
    }
    float s, c;
    sincosf(w * bt, &s, &c);
    const float ex = cx + r_helix * c - x, ey = cy + r_helix * s - y, ez = bt - z;
    const float d = fminf(sqrtf(ex * ex + ey * ey + ez * ez), sqrtf(bd));
    const uint8_t inside = d <= r_tube;
    const uint64_t h = splitmix64(seed ^ splitmix64(uint64_t(i)));
    const float u1 = (float((h >> 40) & 0xFFFFFF) + 1.f) * (1.f / 16777217.f);
    const float u2 = float(h & 0xFFFFFF) * (1.f / 16777216.f);
    const float g = sqrtf(-2.f * logf(u1)) * cosf(6.28318530717958647692f * u2);
    img[i] = float(inside) + sigma * g;
    if (truth) truth[i] = inside;
    if (tangent) {
        // d/dt of (t, cy + Rh sin wt, cx + Rh cos wt) in axis order, unit length
        const float t0 = 1.f, t1 = r_helix * w * c, t2 = -r_helix * w * s;
        const float inv = 1.f / sqrtf(t0 * t0 + t1 * t1 + t2 * t2);
        tangent[3 * i] = t0 * inv;
        tangent[3 * i + 1] = t1 * inv;
        tangent[3 * i + 2] = t2 * inv;
    }
}

}  // namespace

MH_EXPORT int mh_terminal_capacities(mh_ctx *ctx, const float *img, int64_t n, float noise_level,
                                     float eps, float *cs, float *ct, void *stream)
{
    MH_REQUIRE(ctx, img && cs && ct && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    MH_REQUIRE(ctx, noise_level > 0.f && eps > 0.f, "noise level and eps must be positive");
    if (n == 0) return MH_OK;
    const double var = double(noise_level) * noise_level;
    const double norm = 1.0 / std::sqrt(2.0 * M_PI * var);
    const double rq = 1.0 / std::log((norm + eps) / eps);
    hipLaunchKernelGGL(capacities_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, img, n, float(-1.0 / (2.0 * var)), float(norm), eps,
                       float(rq), cs, ct);
    MH_CHECK_LAUNCH(ctx, "capacities_kernel");
    return MH_OK;
}

MH_EXPORT int mh_grad_mag_sq(mh_ctx *ctx, const float *img, const int64_t dims[3], float *out,
                             void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims) && img && out, "bad arguments");
    const int64_t n = dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(grad_mag_sq_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, img, int(dims[0]), int(dims[1]), int(dims[2]), out);
    MH_CHECK_LAUNCH(ctx, "grad_mag_sq_kernel");
    return MH_OK;
}

MH_EXPORT int mh_nonterminal_weight(mh_ctx *ctx, const float *gms, int64_t n, float weighting,
                                    float decay, float *alpha, void *stream)
{
    MH_REQUIRE(ctx, gms && alpha && n >= 0 && n < (int64_t(1) << 31) && decay != 0.f, "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(nonterminal_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, gms, n, weighting, float(1.0 / (double(decay) * decay)),
                       alpha);
    MH_CHECK_LAUNCH(ctx, "nonterminal_kernel");
    return MH_OK;
}

MH_EXPORT int64_t mh_normalize_work_floats(int64_t n)
{
    if (n <= 0) return -1;
    return 2 * int64_t(std::min<int64_t>(1024, mh::blocks_for(n))) + 2;
}

MH_EXPORT int mh_minmax_normalize(mh_ctx *ctx, const float *x, int64_t n, float *out, float *work,
                                  void *stream)
{
    MH_REQUIRE(ctx, x && out && work && n > 0 && n < (int64_t(1) << 31), "bad arguments");
    const unsigned nb = std::min<int64_t>(1024, mh::blocks_for(n));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(minmax_partials_kernel, dim3(nb), dim3(mh::kBlock), 0, s, x, n, work);
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(mh::kBlock), 0, s, work, int(nb));
    hipLaunchKernelGGL(normalize_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, s, x, n,
                       work + 2 * nb, out);
    MH_CHECK_LAUNCH(ctx, "normalize");
    return MH_OK;
}

MH_EXPORT int mh_weighted_dirs(mh_ctx *ctx, const float *dirs3, const float *resp, int64_t n,
                               float *v3, void *stream)
{
    MH_REQUIRE(ctx, dirs3 && resp && v3 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(weighted_dirs_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, dirs3, resp, n, v3);
    MH_CHECK_LAUNCH(ctx, "weighted_dirs_kernel");
    return MH_OK;
}

MH_EXPORT int mh_soa_to_aos3(mh_ctx *ctx, const float *soa, int64_t n, float *aos, void *stream)
{
    MH_REQUIRE(ctx, soa && aos && soa != aos && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, soa, n, aos);
    MH_CHECK_LAUNCH(ctx, "soa_to_aos_kernel");
    return MH_OK;
}

MH_EXPORT int mh_helix_phantom(mh_ctx *ctx, const int64_t dims[3], float r_tube, float r_helix,
                               float pitch, float sigma, uint64_t seed, float *img, uint8_t *truth,
                               float *tangent, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims) && img && pitch > 0.f && r_tube >= 0.f, "bad arguments");
    const int64_t n = dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(helix_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, (hipStream_t)stream,
                       int(dims[0]), int(dims[1]), int(dims[2]), r_tube, r_helix, pitch, sigma, seed,
                       img, truth, tangent);
    MH_CHECK_LAUNCH(ctx, "helix_kernel");
    return MH_OK;
}

MH_EXPORT int64_t mh_initial_u_work_floats(int64_t n)
{
    if (n <= 0) return -1;
    return int64_t(std::min<int64_t>(1024, mh::blocks_for(n))) + 1;
}

MH_EXPORT int mh_initial_u(mh_ctx *ctx, const float *img, int64_t n, float *u, float *work,
                           void *stream)
{
    MH_REQUIRE(ctx, img && u && work && n > 0 && n < (int64_t(1) << 31), "bad arguments");
    const unsigned nb = std::min<int64_t>(1024, mh::blocks_for(n));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reduce_partials_kernel<NanMaxOp>, dim3(nb), dim3(mh::kBlock), 0, s, img, n, work);
    hipLaunchKernelGGL(reduce_final_kernel<NanMaxOp>, dim3(1), dim3(mh::kBlock), 0, s, work, int(nb));
    hipLaunchKernelGGL(initial_u_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, s, img, n,
                       work + nb, u);
    MH_CHECK_LAUNCH(ctx, "initial_u");
    return MH_OK;
}

MH_EXPORT int64_t mh_dice_work_bytes(int64_t n, int nsolves)
{
    if (n <= 0 || nsolves < 1) return -1;
    const int64_t nb = std::min<int64_t>(256, mh::blocks_for(n));
    return int64_t(nsolves) * nb * 3 * int64_t(sizeof(unsigned));
}

MH_EXPORT int mh_dice_counts(mh_ctx *ctx, const float *u, int nsolves, int64_t n,
                             const uint8_t *truth, void *work, int64_t work_bytes,
                             unsigned long long *counts_dev, void *stream)
{
    MH_REQUIRE(ctx, u && truth && work && counts_dev && nsolves >= 1 && nsolves <= 65535 && n > 0 &&
               n < (int64_t(1) << 31), "bad arguments");
    MH_REQUIRE(ctx, work_bytes >= mh_dice_work_bytes(n, nsolves), "workspace too small");
    const unsigned nb = std::min<int64_t>(256, mh::blocks_for(n));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(dice_partials_kernel, dim3(nb, unsigned(nsolves)), dim3(mh::kBlock), 0, s, u, n,
                       truth, static_cast<unsigned *>(work));
    hipLaunchKernelGGL(dice_final_kernel, dim3(unsigned(nsolves)), dim3(64), 0, s,
                       static_cast<const unsigned *>(work), int(nb), counts_dev);
    MH_CHECK_LAUNCH(ctx, "dice counts");
    return MH_OK;
}
