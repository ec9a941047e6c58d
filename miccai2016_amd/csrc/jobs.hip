// Resident max-flow jobs for continuous batching: each job's state lives in
// its own device allocation; one launch sequence advances any set of jobs of
// different sizes by a chunk of iterations. A table of job descriptors maps
// every block of the grid onto its job (binary search over block offsets).
#include <vector>

#include "cmf_voxel.h"
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

constexpr int kCtrlWords = 64;  // [0] stop, [1] number of criteria
constexpr int kMaxJobs = 1024;

struct SumOp {
    __device__ static float identity() { return 0.f; }
    __device__ static float combine(float a, float b) { return a + b; }
};

struct JobDesc {
    int d0, d1, d2, block0;
    float *u;
    const float *alpha, *cs, *ct, *eigs;
    float4 *q4;
    float *hp, *pt, *tmp, *partials, *crit_log;
    int *ctrl;
    float rcc, inv_rcc, rs0, rs1, rs2, ss0, ss1, ss2, sq0, sq1, sq2, rq0, rq1, rq2, error_bound,
        inv_n_rcc;
    mh::AxiParams A;
    int iter0, num_iter, update, maxcheck;
};

int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

int maxcheck_of(const mh_cmf_params *p)
{
    return std::max(1, (p->num_iter + p->update - 1) / std::max(1, p->update));
}

// state: ctrl | criterion log | partials | q4 | hp (3N) | pt | tmp
int64_t state_layout(int64_t n, int maxcheck, char *base, JobDesc *d)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    char *ctrl = take(sizeof(int) * kCtrlWords);
    char *log = take(sizeof(float) * maxcheck);
    char *parts = take(sizeof(float) * mh::blocks_for(n));
    char *q4 = take(sizeof(float4) * n);
    char *hp = take(sizeof(float) * 3 * n);
    char *pt = take(sizeof(float) * n);
    char *tmp = take(sizeof(float) * n);
    if (d) {
        d->ctrl = reinterpret_cast<int *>(ctrl);
        d->crit_log = reinterpret_cast<float *>(log);
        d->partials = reinterpret_cast<float *>(parts);
        d->q4 = reinterpret_cast<float4 *>(q4);
        d->hp = reinterpret_cast<float *>(hp);
        d->pt = reinterpret_cast<float *>(pt);
        d->tmp = reinterpret_cast<float *>(tmp);
    }
    return off;
}

__device__ __forceinline__ int find_job(const JobDesc *__restrict__ J, int nj, unsigned b)
{
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (unsigned(J[mid].block0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(mh::kBlock) jobs_init_kernel(const JobDesc *__restrict__ J, int nj)
{
    const int j = find_job(J, nj, blockIdx.x);
    const JobDesc &D = J[j];
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const int64_t i = int64_t(blockIdx.x - D.block0) * blockDim.x + threadIdx.x;
    if (blockIdx.x == unsigned(D.block0) && threadIdx.x < kCtrlWords) D.ctrl[threadIdx.x] = 0;
    if (i >= n) return;
    const float *e = D.eigs + 3 * i;
    const mh::Quat q = mh::expand_squeezed(e[0], e[1], e[2]);
    D.q4[i] = make_float4(q.w, q.x, q.y, q.z);
    D.hp[i] = 0.f;
    D.hp[n + i] = 0.f;
    D.hp[2 * n + i] = 0.f;
    D.pt[i] = fminf(D.cs[i], D.ct[i]);
    D.tmp[i] = -(D.u[i] * D.rcc);
}

__global__ void __launch_bounds__(mh::kBlock) jobs_p_kernel(const JobDesc *__restrict__ J, int nj, int k)
{
    const int j = find_job(J, nj, blockIdx.x);
    const JobDesc &D = J[j];
    if (D.iter0 + k >= D.num_iter || D.ctrl[0]) return;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const int64_t i = int64_t(blockIdx.x - D.block0) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int i2 = int(i % D.d2), i1 = int((i / D.d2) % D.d1), i0 = int(i / (int64_t(D.d2) * D.d1));
    const int64_t s0 = int64_t(D.d1) * D.d2;
    const float *tmp = D.tmp;
    const float t = tmp[i];
    const float g0 = i0 < D.d0 - 1 ? (tmp[i + s0] - t) * D.ss0 : 0.f;
    const float g1 = i1 < D.d1 - 1 ? (tmp[i + D.d2] - t) * D.ss1 : 0.f;
    const float g2 = i2 < D.d2 - 1 ? (tmp[i + 1] - t) * D.ss2 : 0.f;
    float h0 = D.hp[i], h1 = D.hp[n + i], h2 = D.hp[2 * n + i];
    mh::cmf_p_voxel(h0, h1, h2, g0, g1, g2, D.q4[i], D.sq0, D.sq1, D.sq2, D.rq0, D.rq1, D.rq2, D.A,
                    D.alpha[i]);
    D.hp[i] = h0;
    D.hp[n + i] = h1;
    D.hp[2 * n + i] = h2;
}

__global__ void __launch_bounds__(mh::kBlock) jobs_u_kernel(const JobDesc *__restrict__ J, int nj, int k)
{
    const int j = find_job(J, nj, blockIdx.x);
    const JobDesc &D = J[j];
    const int it = D.iter0 + k;
    if (it >= D.num_iter || D.ctrl[0]) return;  // uniform over the block
    const bool check = (it % D.update) == 0;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const int64_t i = int64_t(blockIdx.x - D.block0) * blockDim.x + threadIdx.x;
    float aerr = 0.f;
    if (i < n) {
        const int i2 = int(i % D.d2), i1 = int((i / D.d2) % D.d1), i0 = int(i / (int64_t(D.d2) * D.d1));
        const int64_t s0 = int64_t(D.d1) * D.d2;
        const float *hp0 = D.hp, *hp1 = D.hp + n, *hp2 = D.hp + 2 * n;
        const float x0 = (i0 < D.d0 - 1 ? hp0[i] : 0.f) - (i0 > 0 ? hp0[i - s0] : 0.f);
        const float x1 = (i1 < D.d1 - 1 ? hp1[i] : 0.f) - (i1 > 0 ? hp1[i - D.d2] : 0.f);
        const float x2 = (i2 < D.d2 - 1 ? hp2[i] : 0.f) - (i2 > 0 ? hp2[i - 1] : 0.f);
        const float div = mh::cmf_div(x0, x1, x2, D.rs0, D.rs1, D.rs2);
        float ui = D.u[i], pti = D.pt[i], ti;
        const float err = mh::cmf_u_voxel(div, pti, ui, ti, D.cs[i], D.ct[i], D.rcc, D.inv_rcc, 1.f, 0.f);
        D.pt[i] = pti;
        D.u[i] = ui;
        D.tmp[i] = ti;
        aerr = fabsf(err);
    }
    if (!check) return;
    aerr = mh::block_reduce<SumOp>(aerr);
    if (threadIdx.x == 0) D.partials[blockIdx.x - D.block0] = aerr;
}

// One work-group per job: on the job's criterion iterations, sum its partials,
// log the criterion and set the stop flag on NaN or convergence.
__global__ void __launch_bounds__(1024) jobs_crit_kernel(const JobDesc *__restrict__ J, int k)
{
    const JobDesc &D = J[blockIdx.x];
    const int it = D.iter0 + k;
    if (it >= D.num_iter || D.ctrl[0] || (it % D.update) != 0) return;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const float s = mh::block_sum_array<SumOp>(D.partials, int(mh::blocks_for(n)));
    if (threadIdx.x == 0) {
        const float crit = s * D.inv_n_rcc;
        const int c = D.ctrl[1];
        if (c < D.maxcheck) D.crit_log[c] = crit;
        D.ctrl[1] = c + 1;
        if (crit != crit) D.ctrl[0] = 1;
        else if (crit < D.error_bound) D.ctrl[0] = 2;
    }
}

__global__ void gather_ctrl_kernel(const JobDesc *__restrict__ J, int nj, int *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    out[2 * j] = J[j].ctrl[0];
    out[2 * j + 1] = J[j].ctrl[1];
}

int build_descs(mh_ctx *ctx, mh_cmf_job *jobs, int nj, std::vector<JobDesc> &D, int64_t &nblocks)
{
    D.resize(nj);
    nblocks = 0;
    for (int j = 0; j < nj; ++j) {
        mh_cmf_job &J = jobs[j];
        MH_REQUIRE(ctx, mh::dims_ok(J.dims), "bad job dims");
        MH_REQUIRE(ctx, J.u_dev && J.eigs_dev && J.alpha_dev && J.cs_dev && J.ct_dev && J.state_dev,
                   "null job pointer");
        MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(J.state_dev) & 255) == 0, "job state must be 256-byte aligned");
        const mh_cmf_params &P = J.params;
        MH_REQUIRE(ctx, P.const_evals == 1 && P.update >= 1 && P.num_iter >= 0 && P.rcc > 0.f, "bad job params");
        const int64_t n = J.dims[0] * J.dims[1] * J.dims[2];
        JobDesc &d = D[j];
        d.d0 = int(J.dims[0]);
        d.d1 = int(J.dims[1]);
        d.d2 = int(J.dims[2]);
        d.block0 = int(nblocks);
        nblocks += mh::blocks_for(n);
        d.u = J.u_dev;
        d.alpha = J.alpha_dev;
        d.cs = J.cs_dev;
        d.ct = J.ct_dev;
        d.eigs = J.eigs_dev;
        state_layout(n, maxcheck_of(&P), static_cast<char *>(J.state_dev), &d);
        d.rcc = P.rcc;
        d.rs0 = P.rs[0]; d.rs1 = P.rs[1]; d.rs2 = P.rs[2];
        d.ss0 = P.ss[0]; d.ss1 = P.ss[1]; d.ss2 = P.ss[2];
        d.sq0 = P.sqev[0]; d.sq1 = P.sqev[1]; d.sq2 = P.sqev[2];
        d.rq0 = float(1.0 / double(P.sqev[0]));
        d.rq1 = float(1.0 / double(P.sqev[1]));
        d.rq2 = float(1.0 / double(P.sqev[2]));
        d.A = mh::axi_params(&P);
        d.inv_rcc = float(1.0 / double(P.rcc));
        d.error_bound = P.error_bound;
        d.inv_n_rcc = float(1.0 / (double(n) * double(P.rcc)));
        d.iter0 = J.iter;
        d.num_iter = P.num_iter;
        d.update = P.update;
        d.maxcheck = maxcheck_of(&P);
    }
    MH_REQUIRE(ctx, nblocks < (int64_t(1) << 31), "too many blocks");
    return MH_OK;
}

int upload_descs(mh_ctx *ctx, const std::vector<JobDesc> &D, hipStream_t s)
{
    const size_t need = sizeof(JobDesc) * D.size() + sizeof(int) * 2 * D.size() + 256;
    if (ctx->desc_bytes < need) {
        if (ctx->desc) MH_CHECK(ctx, hipFree(ctx->desc));
        ctx->desc = nullptr;
        ctx->desc_bytes = 0;
        MH_CHECK(ctx, hipMalloc(&ctx->desc, need));
        ctx->desc_bytes = need;
    }
    // synchronous with respect to the host: D may go out of scope afterwards
    MH_CHECK(ctx, hipMemcpyAsync(ctx->desc, D.data(), sizeof(JobDesc) * D.size(), hipMemcpyHostToDevice, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    return MH_OK;
}

}  // namespace

MH_EXPORT int64_t mh_cmf_job_state_bytes(int64_t n, const mh_cmf_params *params)
{
    if (n <= 0 || n >= (int64_t(1) << 31) || !params || params->update < 1) return -1;
    return state_layout(n, maxcheck_of(params), nullptr, nullptr);
}

MH_EXPORT int mh_cmf_jobs_init(mh_ctx *ctx, mh_cmf_job *jobs, int nj, void *stream)
{
    MH_REQUIRE(ctx, jobs && nj >= 1 && nj <= kMaxJobs, "1..1024 jobs");
    std::vector<JobDesc> D;
    int64_t nblocks = 0;
    int rc = build_descs(ctx, jobs, nj, D, nblocks);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = upload_descs(ctx, D, s);
    if (rc) return rc;
    hipLaunchKernelGGL(jobs_init_kernel, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s,
                       static_cast<const JobDesc *>(ctx->desc), nj);
    MH_CHECK_LAUNCH(ctx, "jobs_init_kernel");
    for (int j = 0; j < nj; ++j) {
        jobs[j].iter = 0;
        jobs[j].stopped = 0;
    }
    return MH_OK;
}

MH_EXPORT int mh_cmf_jobs_step(mh_ctx *ctx, mh_cmf_job *jobs, int nj, int iters, void *stream)
{
    MH_REQUIRE(ctx, jobs && nj >= 1 && nj <= kMaxJobs && iters >= 0, "1..1024 jobs, iters >= 0");
    std::vector<JobDesc> D;
    int64_t nblocks = 0;
    int rc = build_descs(ctx, jobs, nj, D, nblocks);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = upload_descs(ctx, D, s);
    if (rc) return rc;
    const JobDesc *dd = static_cast<const JobDesc *>(ctx->desc);
    int most = 0;
    for (int j = 0; j < nj; ++j)
        if (!jobs[j].stopped) most = std::max(most, std::min(iters, D[j].num_iter - D[j].iter0));
    for (int k = 0; k < most; ++k) {
        hipLaunchKernelGGL(jobs_p_kernel, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, dd, nj, k);
        hipLaunchKernelGGL(jobs_u_kernel, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, dd, nj, k);
        bool any_check = false;
        for (int j = 0; j < nj && !any_check; ++j) {
            const int it = D[j].iter0 + k;
            any_check = !jobs[j].stopped && it < D[j].num_iter && (it % D[j].update) == 0;
        }
        if (any_check) hipLaunchKernelGGL(jobs_crit_kernel, dim3(nj), dim3(1024), 0, s, dd, k);
        MH_CHECK_LAUNCH(ctx, "job kernels");
    }
    int *out = reinterpret_cast<int *>(static_cast<char *>(ctx->desc) + sizeof(JobDesc) * nj);
    hipLaunchKernelGGL(gather_ctrl_kernel, dim3((nj + 255) / 256), dim3(256), 0, s, dd, nj, out);
    std::vector<int> ctrl(2 * nj);
    MH_CHECK(ctx, hipMemcpyAsync(ctrl.data(), out, sizeof(int) * 2 * nj, hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    for (int j = 0; j < nj; ++j) {
        mh_cmf_job &J = jobs[j];
        J.stopped = ctrl[2 * j];
        if (J.stopped) J.iter = (ctrl[2 * j + 1] - 1) * D[j].update + 1;
        else J.iter = std::min(D[j].num_iter, D[j].iter0 + iters);
    }
    return MH_OK;
}
