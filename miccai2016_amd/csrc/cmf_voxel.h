// Per-voxel arithmetic of one ATV continuous max-flow iteration, shared by
// every solver variant (single, batched, ragged, resident jobs) so that all of
// them compute bit-identical iterations.
//
// H = R(q) diag(sqrt(lambda)). The p-step takes the forward differences of the
// residual tmp, scaled by ss, and the previous hp = H p:
//   p = H^-1 hp + H^T g, projected onto |p| <= alpha, hp' = H p.
// The u-step takes div = -(adjoint of the forward difference) of hp':
//   ps = min(div + pt - rcc (u - ub), cs), pt' = min(ps - div + rcc (u - uf), ct),
//   err = div - ps + pt', u' = u - err / rcc, tmp' = err - rcc u'
// with min(x, y) = (y < x ? y : x) as in the source algorithm's kernel
// language (a NaN residual propagates into the criterion, a NaN capacity is
// dropped) and ub = 1, uf = 0 for the unsupervised cut.
#pragma once

#include "mh_common.h"

namespace mh {

__device__ __forceinline__ float kmin(float x, float y) { return y < x ? y : x; }

// Every helper below is evaluated without implicit contraction into fused
// multiply-adds: the fused operations are the explicit fmaf calls, so that
// each solver variant, whatever the surrounding code, rounds identically.

// divergence from the backward differences x_k of hp_k, scaled by rs
__device__ __forceinline__ float cmf_div(float x0, float x1, float x2, float rs0, float rs1,
                                         float rs2)
{
#pragma clang fp contract(off)
    return fmaf(x2, rs2, fmaf(x1, rs1, x0 * rs0));
}

// sq: sqrt of the eigenvalues of the frame; rq = 1 / sq (formed once on the
// host, in double). The norm uses the hardware square root and the
// projection the hardware reciprocal, as the reference's kernel uses
// native_sqrt (update_qp.cl:35, 215-221).
__device__ __forceinline__ void cmf_p_voxel_m(float &h0, float &h1, float &h2, float g0, float g1,
                                              float g2, const Mat3 &R, float sq0, float sq1,
                                              float sq2, float rq0, float rq1, float rq2,
                                              float alpha)
{
#pragma clang fp contract(off)
    const float *m = R.m;
    // R^T hp and R^T g
    const float a0 = fmaf(m[6], h2, fmaf(m[3], h1, m[0] * h0));
    const float a1 = fmaf(m[7], h2, fmaf(m[4], h1, m[1] * h0));
    const float a2 = fmaf(m[8], h2, fmaf(m[5], h1, m[2] * h0));
    const float b0 = fmaf(m[6], g2, fmaf(m[3], g1, m[0] * g0));
    const float b1 = fmaf(m[7], g2, fmaf(m[4], g1, m[1] * g0));
    const float b2 = fmaf(m[8], g2, fmaf(m[5], g1, m[2] * g0));
    float p0 = fmaf(a0, rq0, sq0 * b0);
    float p1 = fmaf(a1, rq1, sq1 * b1);
    float p2 = fmaf(a2, rq2, sq2 * b2);
    const float nrm = __builtin_amdgcn_sqrtf(fmaf(p2, p2, fmaf(p1, p1, p0 * p0)));
    const float sc = nrm > alpha ? alpha * __builtin_amdgcn_rcpf(nrm) : 1.f;
    p0 *= sc * sq0;
    p1 *= sc * sq1;
    p2 *= sc * sq2;
    h0 = fmaf(m[2], p2, fmaf(m[1], p1, m[0] * p0));
    h1 = fmaf(m[5], p2, fmaf(m[4], p1, m[3] * p0));
    h2 = fmaf(m[8], p2, fmaf(m[7], p1, m[6] * p0));
}

// the same with the rotation given as the unit quaternion (w, x, y, z) in qq
__device__ __forceinline__ void cmf_p_voxel(float &h0, float &h1, float &h2, float g0, float g1,
                                            float g2, const float4 qq, float sq0, float sq1,
                                            float sq2, float rq0, float rq1, float rq2, float alpha)
{
    cmf_p_voxel_m(h0, h1, h2, g0, g1, g2, quat_to_mat(Quat{qq.x, qq.y, qq.z, qq.w}), sq0, sq1, sq2,
                  rq0, rq1, rq2, alpha);
}

// Returns err; updates pt, u and writes the next residual to tmp. inv_rcc =
// 1 / rcc (formed once on the host, in double).
__device__ __forceinline__ float cmf_u_voxel(float div, float &pt, float &u, float &tmp, float cs,
                                             float ct, float rcc, float inv_rcc, float ub, float uf)
{
#pragma clang fp contract(off)
    const float ps = kmin(fmaf(-rcc, u - ub, div + pt), cs);
    pt = kmin(fmaf(rcc, u - uf, ps - div), ct);
    const float err = div - ps + pt;
    u = u - err * inv_rcc;
    tmp = fmaf(-u, rcc, err);
    return err;
}

}  // namespace mh
