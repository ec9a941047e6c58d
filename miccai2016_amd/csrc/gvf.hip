// Gradient vector flow on gfx950: one fused explicit-Euler step per launch,
// ping-ponging the three SoA components, mirrored neighbours at the faces.
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

struct SumOp {
    __device__ static float identity() { return 0.f; }
    __device__ static float combine(float a, float b) { return a + b; }
};

// g' = g - dt * (r (g - v) - mu * lap g), lap with the 7-point stencil and
// the neighbour across a face replaced by the one on the inner side.
// With `partials`, each block stores the sum of |delta| over its voxels.
__global__ void __launch_bounds__(mh::kBlock)
gvf_step_kernel(float *__restrict__ out, const float *__restrict__ g, const float *__restrict__ v,
                const float *__restrict__ r, int d0, int d1, int d2, float mu, float dt, float rss0,
                float rss1, float rss2, float *__restrict__ partials)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    float dmag = 0.f;
    if (i < n) {
        const int i2 = int(i % d2);
        const int i1 = int((i / d2) % d1);
        const int i0 = int(i / (int64_t(d2) * d1));
        const int64_t s0 = int64_t(d1) * d2, s1 = d2;
        const int64_t b0 = i0 > 0 ? -s0 : s0, a0 = i0 < d0 - 1 ? s0 : -s0;
        const int64_t b1 = i1 > 0 ? -s1 : s1, a1 = i1 < d1 - 1 ? s1 : -s1;
        const int64_t b2 = i2 > 0 ? -1 : 1, a2 = i2 < d2 - 1 ? 1 : -1;
        const float ri = r[i];
        float dsq = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *gc = g + c * n;
            const float gi = gc[i];
            float lap = (gc[i + b0] + gc[i + a0] - 2.f * gi) * rss0;
            lap += (gc[i + b1] + gc[i + a1] - 2.f * gi) * rss1;
            lap += (gc[i + b2] + gc[i + a2] - 2.f * gi) * rss2;
            const float delta = fmaf(-mu, lap, ri * (gi - v[c * n + i]));
            out[c * n + i] = fmaf(-dt, delta, gi);
            dsq += delta * delta;
        }
        dmag = sqrtf(dsq);
    }
    if (!partials) return;
    dmag = mh::block_reduce<SumOp>(dmag);
    if (threadIdx.x == 0) partials[blockIdx.x] = dmag;
}

// Sum of the per-block partials in block order (one work-group).
__global__ void __launch_bounds__(mh::kBlock)
sum_partials_kernel(const float *__restrict__ partials, int nparts, float *__restrict__ out)
{
    float s = 0.f;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) s += partials[k];
    s = mh::block_reduce<SumOp>(s);
    if (threadIdx.x == 0) *out = s;
}

__global__ void __launch_bounds__(mh::kBlock)
response_kernel(float *__restrict__ r, const float *__restrict__ v, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = v[i], b = v[n + i], c = v[2 * n + i];
    r[i] = a * a + b * b + c * c;
}

__global__ void __launch_bounds__(mh::kBlock) normalize_kernel(float *g, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = g[i], b = g[n + i], c = g[2 * n + i];
    const float nrm = sqrtf(a * a + b * b + c * c);
    if (nrm != 0.f) {
        const float k = 1.f / nrm;
        g[i] = k * a;
        g[n + i] = k * b;
        g[2 * n + i] = k * c;
    }
}

}  // namespace

MH_EXPORT int64_t mh_gvf_work_floats(const int64_t dims[3])
{
    if (!mh::dims_ok(dims)) return -1;
    return int64_t(mh::blocks_for(dims[0] * dims[1] * dims[2])) + 64;
}

MH_EXPORT int mh_gvf_response(mh_ctx *ctx, float *r, const float *v3, int64_t n, void *stream)
{
    MH_REQUIRE(ctx, r && v3 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(response_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, r, v3, n);
    MH_CHECK_LAUNCH(ctx, "response_kernel");
    return MH_OK;
}

MH_EXPORT int mh_gvf(mh_ctx *ctx, float *g3, float *swap3, const float *v3, const float *r,
                     const int64_t dims[3], const mh_gvf_params *p, int iters, float *mag,
                     float *work, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, dims[0] >= 2 && dims[1] >= 2 && dims[2] >= 2, "every dim must be >= 2");
    MH_REQUIRE(ctx, g3 && swap3 && v3 && r && p && iters >= 0, "bad arguments");
    MH_REQUIRE(ctx, !mag || work, "mag_dev needs work_dev");
    const int64_t n = dims[0] * dims[1] * dims[2];
    const unsigned nb = mh::blocks_for(n);
    hipStream_t s = (hipStream_t)stream;
    float *src = g3, *dst = swap3;
    for (int it = 0; it < iters; ++it) {
        float *parts = (mag && it == iters - 1) ? work : nullptr;
        hipLaunchKernelGGL(gvf_step_kernel, dim3(nb), dim3(mh::kBlock), 0, s, dst, src, v3, r,
                           int(dims[0]), int(dims[1]), int(dims[2]), p->mu, p->dt, p->rss[0],
                           p->rss[1], p->rss[2], parts);
        MH_CHECK_LAUNCH(ctx, "gvf_step_kernel");
        if (parts) {
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(mh::kBlock), 0, s, parts, int(nb), mag);
            MH_CHECK_LAUNCH(ctx, "sum_partials_kernel");
        }
        float *t = src;
        src = dst;
        dst = t;
    }
    if (src != g3) MH_CHECK(ctx, hipMemcpyAsync(g3, src, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, s));
    return MH_OK;
}

MH_EXPORT int mh_gvf_normalize(mh_ctx *ctx, float *g3, int64_t n, void *stream)
{
    MH_REQUIRE(ctx, g3 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(normalize_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, g3, n);
    MH_CHECK_LAUNCH(ctx, "normalize_kernel");
    return MH_OK;
}
