// Gradient vector flow on gfx950: explicit-Euler steps on the three SoA
// components, ping-ponging two buffers, mirrored neighbours at the faces;
// one step per launch, or two per launch in a sweep along axis 0.
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

struct SumOp {
    __device__ static float identity() { return 0.f; }
    __device__ static float combine(float a, float b) { return a + b; }
};

// One step of one component at one voxel: out = g - dt (r (g - v) - mu lap),
// lap = sum_k rss_k ((g[-e_k] + g[+e_k]) - 2 g) over the axes, with the
// neighbour across a face replaced by the one on the inner side (m_k: the
// lower neighbour, p_k: the upper). Only explicit fused operations, so that
// every kernel below rounds identically. Returns delta.
__device__ __forceinline__ float gvf_voxel(float gi, float m0, float p0, float m1, float p1,
                                           float m2, float p2, float v, float r, float mu,
                                           float dt, float rss0, float rss1, float rss2,
                                           float &out)
{
#pragma clang fp contract(off)
    const float t0 = fmaf(-2.f, gi, m0 + p0);
    const float t1 = fmaf(-2.f, gi, m1 + p1);
    const float t2 = fmaf(-2.f, gi, m2 + p2);
    const float lap = fmaf(t2, rss2, fmaf(t1, rss1, t0 * rss0));
    const float delta = fmaf(-mu, lap, r * (gi - v));
    out = fmaf(-dt, delta, gi);
    return delta;
}

// |delta| of a voxel from its three components
__device__ __forceinline__ float gvf_mag(float d0, float d1, float d2)
{
#pragma clang fp contract(off)
    return sqrtf(fmaf(d2, d2, fmaf(d1, d1, d0 * d0)));
}

// One step over the volume. With `partials`, each block stores the sum of
// |delta| over its voxels.
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
        float dl[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *gc = g + c * n;
            float o;
            dl[c] = gvf_voxel(gc[i], gc[i + b0], gc[i + a0], gc[i + b1], gc[i + a1], gc[i + b2],
                              gc[i + a2], v[c * n + i], ri, mu, dt, rss0, rss1, rss2, o);
            out[c * n + i] = o;
        }
        dmag = gvf_mag(dl[0], dl[1], dl[2]);
    }
    if (!partials) return;
    dmag = mh::block_reduce<SumOp>(dmag);
    if (threadIdx.x == 0) partials[blockIdx.x] = dmag;
}

// Two steps per launch, temporally blocked along axis 0 (as the max-flow
// sweep is). A work-group owns an output tile of TY x TX voxels in the
// (axis 1, axis 2) plane and a chunk of planes [z0, z1), one thread per point
// of the tile widened by 2 on every side (E2). Step z loads g of plane z + 2
// ahead of its use, puts g of plane z into LDS, and after the step's one
// barrier computes
//   level A: the first step at plane z on the tile widened by 1 (E1), from
//            LDS (in-plane neighbours) and registers (planes z - 1, z + 1),
//            into LDS by plane parity;
//   level B: the second step at plane z - 1 on the tile, from the level-A
//            results of the previous step in LDS (written before this
//            step's barrier) and in registers (planes z - 2, z - 1, z),
//            and stores it.
// LDS holds two planes of each (by plane parity), so one barrier per step
// orders every write before the reads that need it and after the reads of
// the plane two steps back. The per-voxel arithmetic is gvf_voxel's, so the
// result equals two launches of gvf_step_kernel bit for bit. With
// `partials`, each work-group stores the sum of |delta| of its second step.
template <int TY, int TX, int NT>
__global__ void __launch_bounds__(NT)
gvf_sweep_kernel(float *__restrict__ out, const float *__restrict__ g, const float *__restrict__ v,
                 const float *__restrict__ r, int d0, int d1, int d2, float mu, float dt, float rss0,
                 float rss1, float rss2, int zchunk, float *__restrict__ partials)
{
    constexpr int EY = TY + 4, EX = TX + 4, NE = EY * EX;
    static_assert(NE <= NT, "one point per thread");
    __shared__ float sG[2][3][NE];  // g of plane z, by plane parity
    __shared__ float sH[2][3][NE];  // level-A result of plane z, by plane parity
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int ntx = (d2 + TX - 1) / TX, nty = (d1 + TY - 1) / TY;
    const int tx = int(b % unsigned(ntx)), ty = int((b / unsigned(ntx)) % unsigned(nty));
    const int tz = int(b / unsigned(ntx * nty));
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * zchunk;
    const int z1 = min(z0 + zchunk, d0);
    const int t = int(threadIdx.x);
    const bool live = t < NE;
    const int yy = live ? t / EX : 0, xx = live ? t % EX : 0;
    const int p = yy * EX + xx;
    const int y = y0 - 2 + yy, x = x0 - 2 + xx;
    const bool inv = live && y >= 0 && y < d1 && x >= 0 && x < d2;
    const bool inE1 = inv && yy >= 1 && yy <= TY + 2 && xx >= 1 && xx <= TX + 2;
    const bool inO = inv && yy >= 2 && yy <= TY + 1 && xx >= 2 && xx <= TX + 1;
    // in-plane neighbours in LDS, mirrored at the faces (inside E2 for E1)
    const int pym = y > 0 ? p - EX : p + EX, pyp = y < d1 - 1 ? p + EX : p - EX;
    const int pxm = x > 0 ? p - 1 : p + 1, pxp = x < d2 - 1 ? p + 1 : p - 1;
    const int64_t n = int64_t(d0) * d1 * d2, st0 = int64_t(d1) * d2;
    const int64_t ip = inv ? int64_t(y) * d2 + x : 0;
    auto at = [&](int z) { return int64_t(z) * st0 + ip; };

    const int zs = max(z0 - 1, 0);  // first level-A plane
    const int zend = z1;            // last step (level A only below d0)
    float gm[3], gc[3], gp[3], gn[3];    // g at planes z - 1, z, z + 1 (mirrored), z + 2
    float vA[3], vB[3], vN[3];           // v at planes z, z - 1, z + 1
    float rA = 0.f, rB = 0.f, rN = 0.f;
    float h2[3], h1[3], h0[3];           // level A at planes z - 2, z - 1, z
    float dsum = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gm[c] = gc[c] = gp[c] = gn[c] = vA[c] = vB[c] = vN[c] = h2[c] = h1[c] = h0[c] = 0.f;
    }
    if (inv) {
        const int zm = zs > 0 ? zs - 1 : zs + 1, zp = zs < d0 - 1 ? zs + 1 : zs - 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gc[c] = g[c * n + at(zs)];
            gm[c] = g[c * n + at(zm)];
            gp[c] = g[c * n + at(zp)];
        }
    }
    if (inE1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) vA[c] = v[c * n + at(zs)];
        rA = r[at(zs)];
    }
    for (int z = zs; z <= zend; ++z) {
        const bool la = z < d0;                  // level A at plane z
        const bool lb = z - 1 >= z0;             // level B at plane z - 1 (< z1 since z <= z1)
        const int q = z & 1, qb = (z - 1) & 1;
        // inputs of the next step
        if (inv && z + 1 < d0) {
            if (z + 2 < d0) {
#pragma unroll
                for (int c = 0; c < 3; ++c) gn[c] = g[c * n + at(z + 2)];
            } else {  // plane z + 1 is the last: its upper neighbour mirrors to plane z
#pragma unroll
                for (int c = 0; c < 3; ++c) gn[c] = gc[c];
            }
        }
        if (inE1 && z + 1 < d0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) vN[c] = v[c * n + at(z + 1)];
            rN = r[at(z + 1)];
        }
        if (la && live) {
#pragma unroll
            for (int c = 0; c < 3; ++c) sG[q][c][p] = gc[c];
        }
        __syncthreads();
        if (la && inE1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                h0[c] = 0.f;
                (void)gvf_voxel(gc[c], gm[c], gp[c], sG[q][c][pym], sG[q][c][pyp], sG[q][c][pxm],
                                sG[q][c][pxp], vA[c], rA, mu, dt, rss0, rss1, rss2, h0[c]);
                sH[q][c][p] = h0[c];
            }
        }
        if (lb && inO) {
            const int zb = z - 1;
            float dl[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float below = zb > 0 ? h2[c] : h0[c];
                const float above = zb < d0 - 1 ? h0[c] : h2[c];
                float o;
                dl[c] = gvf_voxel(h1[c], below, above, sH[qb][c][pym], sH[qb][c][pyp],
                                  sH[qb][c][pxm], sH[qb][c][pxp], vB[c], rB, mu, dt, rss0, rss1,
                                  rss2, o);
                out[c * n + at(zb)] = o;
            }
            if (partials) dsum += gvf_mag(dl[0], dl[1], dl[2]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gm[c] = gc[c];
            gc[c] = gp[c];
            gp[c] = gn[c];
            h2[c] = h1[c];
            h1[c] = h0[c];
            vB[c] = vA[c];
            vA[c] = vN[c];
        }
        rB = rA;
        rA = rN;
    }
    if (!partials) return;
    dsum = mh::block_reduce<SumOp>(dsum);
    if (threadIdx.x == 0) partials[blockIdx.x] = dsum;
}

constexpr int kGvfTY = 11, kGvfTX = 64, kGvfNT = 1024;

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

// Plane chunk of the two-step kernel from the launch cost model of the
// max-flow sweep: work-groups run in ceil(groups / slots) waves of about
// (chunk + 2) plane steps; the chunk must leave at most max_groups groups.
int gvf_zchunk(mh_ctx *ctx, const int64_t dims[3], int64_t max_groups)
{
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
        cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&gvf_sweep_kernel<kGvfTY, kGvfTX, kGvfNT>), kGvfNT,
            0) != hipSuccess)
        per = 0;
    (void)hipGetLastError();
    const int64_t slots = (cus > 0 && per > 0) ? int64_t(cus) * per : 256;
    const int64_t tiles = ((dims[2] + kGvfTX - 1) / kGvfTX) * ((dims[1] + kGvfTY - 1) / kGvfTY);
    const int d0 = int(dims[0]);
    int best = 0;
    int64_t best_cost = -1;
    for (int z = 1; z <= d0; ++z) {
        const int64_t groups = tiles * ((d0 + z - 1) / z);
        if (groups > max_groups) continue;
        const int64_t cost = ((groups + slots - 1) / slots) * (z + 2);
        if (best_cost < 0 || cost < best_cost || (cost == best_cost && z > best)) {
            best_cost = cost;
            best = z;
        }
    }
    return best;  // 0: none fits
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
    // two steps per launch where the volume fills the device with them
    // (automatic: from 2^20 voxels on); the partial sums of the last launch
    // fit in the workspace's nb floats
    int zc = 0;
    if (ctx->gvf_sweep != 1 && iters >= 2 && (ctx->gvf_sweep == 2 || n >= (int64_t(1) << 20)))
        zc = gvf_zchunk(ctx, dims, int64_t(nb));
    const int64_t groups = zc ? ((dims[2] + kGvfTX - 1) / kGvfTX) * ((dims[1] + kGvfTY - 1) / kGvfTY) *
                                    ((dims[0] + zc - 1) / zc)
                              : 0;
    int it = 0;
    auto step = [&](float *parts) {
        hipLaunchKernelGGL(gvf_step_kernel, dim3(nb), dim3(mh::kBlock), 0, s, dst, src, v3, r,
                           int(dims[0]), int(dims[1]), int(dims[2]), p->mu, p->dt, p->rss[0],
                           p->rss[1], p->rss[2], parts);
        if (parts)
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(mh::kBlock), 0, s, parts, int(nb), mag);
        float *t = src;
        src = dst;
        dst = t;
        ++it;
    };
    if (zc && iters % 2 == 1) step((mag && iters == 1) ? work : nullptr);  // an odd first step
    while (zc && it + 2 <= iters) {
        float *parts = (mag && it + 2 == iters) ? work : nullptr;
        hipLaunchKernelGGL((gvf_sweep_kernel<kGvfTY, kGvfTX, kGvfNT>), dim3(unsigned(groups)),
                           dim3(kGvfNT), 0, s, dst, src, v3, r, int(dims[0]), int(dims[1]),
                           int(dims[2]), p->mu, p->dt, p->rss[0], p->rss[1], p->rss[2], zc, parts);
        if (parts)
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(mh::kBlock), 0, s, parts,
                               int(groups), mag);
        float *t = src;
        src = dst;
        dst = t;
        it += 2;
    }
    while (it < iters) step((mag && it == iters - 1) ? work : nullptr);
    MH_CHECK_LAUNCH(ctx, "gvf kernels");
    if (src != g3) MH_CHECK(ctx, hipMemcpyAsync(g3, src, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, s));
    return MH_OK;
}

MH_EXPORT int mh_gvf_set_sweep(mh_ctx *ctx, int mode)
{
    if (!ctx) return MH_ERR_ARG;
    MH_REQUIRE(ctx, mode >= 0 && mode <= 2, "mode 0 (automatic), 1 (off) or 2 (on)");
    ctx->gvf_sweep = mode;
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
