// Library context, fill and whole-array reductions.
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

constexpr size_t kScratchBytes = 64 * 1024;
constexpr int kReduceBlocks = 1024;

__global__ void __launch_bounds__(mh::kBlock) fill_kernel(float *__restrict__ x, int64_t n, float v)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

struct SumAbsOp {
    __device__ static float identity() { return 0.f; }
    __device__ static float load(float v) { return fabsf(v); }
    __device__ static float combine(float a, float b) { return a + b; }
};

struct MaxOp {  // NaN-neutral maximum (fmax semantics)
    __device__ static float identity() { return __builtin_nanf(""); }
    __device__ static float load(float v) { return v; }
    __device__ static float combine(float a, float b) { return fmaxf(a, b); }
};

// Grid-stride partials per block, then one work-group combines them in block
// order, so the result does not depend on scheduling.
template <class Op>
__global__ void __launch_bounds__(mh::kBlock)
reduce_kernel(const float *__restrict__ x, int64_t n, float *__restrict__ partials)
{
    float acc = Op::identity();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        acc = Op::combine(acc, Op::load(x[i]));
    acc = mh::block_reduce<Op>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <class Op>
__global__ void __launch_bounds__(mh::kBlock)
combine_kernel(const float *__restrict__ partials, int nparts, float *__restrict__ out)
{
    float r = Op::identity();
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) r = Op::combine(r, partials[b]);
    r = mh::block_reduce<Op>(r);
    if (threadIdx.x == 0) *out = r;
}

template <class Op>
int reduce_impl(mh_ctx *ctx, const float *x, int64_t n, float *out, void *stream)
{
    MH_REQUIRE(ctx, ctx && x && out && n > 0, "null pointer or empty input");
    const unsigned nb = std::min<int64_t>(kReduceBlocks, mh::blocks_for(n));
    float *partials = reinterpret_cast<float *>(static_cast<unsigned char *>(ctx->scratch) + 4096);
    hipLaunchKernelGGL(reduce_kernel<Op>, dim3(nb), dim3(mh::kBlock), 0, (hipStream_t)stream, x, n,
                       partials);
    hipLaunchKernelGGL(combine_kernel<Op>, dim3(1), dim3(mh::kBlock), 0, (hipStream_t)stream, partials,
                       int(nb), out);
    MH_CHECK_LAUNCH(ctx, "reduce_kernel");
    return MH_OK;
}

}  // namespace

MH_EXPORT int mh_version(void) { return 100; }

MH_EXPORT int mh_create(int device, mh_ctx **out)
{
    if (!out) return MH_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return MH_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return MH_ERR_HIP;
    mh_ctx *ctx = new mh_ctx();
    ctx->device = device;
    if (hipMalloc(&ctx->scratch, kScratchBytes) != hipSuccess) {
        delete ctx;
        return MH_ERR_HIP;
    }
    ctx->scratch_bytes = kScratchBytes;
    // the counters at the head of the block must start at zero; this runs once
    // at creation, outside any capture
    if (hipMemset(ctx->scratch, 0, kScratchBytes) != hipSuccess) {
        (void)hipFree(ctx->scratch);
        delete ctx;
        return MH_ERR_HIP;
    }
    // the no-allocation solver's array on this device; a failure is kept and
    // reported by that entry point (every other entry point still works)
    if (mh_noalloc_setup(ctx) != MH_OK) ctx->noalloc_error = ctx->last_error;
    *out = ctx;
    return MH_OK;
}

MH_EXPORT int mh_destroy(mh_ctx *ctx)
{
    if (!ctx) return MH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->desc) (void)hipFree(ctx->desc);
    if (ctx->graph_stream) (void)hipStreamDestroy(ctx->graph_stream);
    delete ctx;
    return MH_OK;
}

MH_EXPORT const char *mh_last_error(mh_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "null context";
}

MH_EXPORT int mh_fill(mh_ctx *ctx, float *x, int64_t n, float value, void *stream)
{
    MH_REQUIRE(ctx, x && n >= 0, "null pointer or negative size");
    if (n == 0) return MH_OK;
    const unsigned nb = std::min<int64_t>(65535, mh::blocks_for(n));
    hipLaunchKernelGGL(fill_kernel, dim3(nb), dim3(mh::kBlock), 0, (hipStream_t)stream, x, n, value);
    MH_CHECK_LAUNCH(ctx, "fill_kernel");
    return MH_OK;
}

MH_EXPORT int mh_reduce_sum_abs(mh_ctx *ctx, const float *x, int64_t n, float *out, void *stream)
{
    return reduce_impl<SumAbsOp>(ctx, x, n, out, stream);
}

MH_EXPORT int mh_reduce_max(mh_ctx *ctx, const float *x, int64_t n, float *out, void *stream)
{
    return reduce_impl<MaxOp>(ctx, x, n, out, stream);
}
