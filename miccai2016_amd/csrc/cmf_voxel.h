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

// Constant eigenvalues (l0, l1, l1), the two smaller ones equal (the
// experiment's ATV (1, a, a) and TV (1, 1, 1)): then H H^T = A =
// l1 I + (l0 - l1) v v^T with v = R e0, the frame's first column, and the
// p-step reduces to
//   z = hp + A g,  |p|^2 = z^T A^-1 z = (v.z)^2 / l0 + |z - (v.z) v|^2 / l1,
//   hp' = min(1, alpha / |p|) z,
// the same map as cmf_p_voxel_m (p = H^-1 hp + H^T g, hp' = H proj(p)) in
// about 30 operations instead of 47, needing v only. Both terms of |p|^2 are
// non-negative, so nothing cancels.
struct AxiParams {
    int on;       // the eigenvalues are constant and l1 == l2
    float l0, l1;
    float dl;     // l0 - l1
    float il0, il1;
};

// AxiParams of a solve (host): on when the eigenvalues are constant and the
// two smaller square roots are equal; l = sq^2, formed in double
inline AxiParams axi_params(const mh_cmf_params *p)
{
    AxiParams A{};
    A.on = p->const_evals && p->sqev[1] == p->sqev[2];
    const double l0 = double(p->sqev[0]) * double(p->sqev[0]);
    const double l1 = double(p->sqev[1]) * double(p->sqev[1]);
    A.l0 = float(l0);
    A.l1 = float(l1);
    A.dl = float(l0 - l1);
    A.il0 = l0 > 0 ? float(1.0 / l0) : 0.f;
    A.il1 = l1 > 0 ? float(1.0 / l1) : 0.f;
    return A;
}

// v = first column of the rotation of the unit quaternion qq = (w, x, y, z)
__device__ __forceinline__ void frame_axis(const float4 qq, float &v0, float &v1, float &v2)
{
#pragma clang fp contract(off)
    const float w = qq.x, x = qq.y, y = qq.z, z = qq.w;
    v0 = 1.f - 2.f * (y * y + z * z);
    v1 = 2.f * (x * y + w * z);
    v2 = 2.f * (x * z - w * y);
}

__device__ __forceinline__ void cmf_p_voxel_axi(float &h0, float &h1, float &h2, float g0, float g1,
                                                float g2, float v0, float v1, float v2,
                                                const AxiParams &A, float alpha)
{
#pragma clang fp contract(off)
    const float t = A.dl * fmaf(v2, g2, fmaf(v1, g1, v0 * g0));
    const float z0 = fmaf(t, v0, fmaf(A.l1, g0, h0));
    const float z1 = fmaf(t, v1, fmaf(A.l1, g1, h1));
    const float z2 = fmaf(t, v2, fmaf(A.l1, g2, h2));
    const float c = fmaf(v2, z2, fmaf(v1, z1, v0 * z0));
    const float e0 = fmaf(-c, v0, z0), e1 = fmaf(-c, v1, z1), e2 = fmaf(-c, v2, z2);
    const float nrm = __builtin_amdgcn_sqrtf(
        fmaf(A.il1, fmaf(e2, e2, fmaf(e1, e1, e0 * e0)), A.il0 * (c * c)));
    const float sc = nrm > alpha ? alpha * __builtin_amdgcn_rcpf(nrm) : 1.f;
    h0 = sc * z0;
    h1 = sc * z1;
    h2 = sc * z2;
}

// the p-step with the rotation given as the unit quaternion (w, x, y, z) in
// qq: the axial form when A.on, else the general one
__device__ __forceinline__ void cmf_p_voxel(float &h0, float &h1, float &h2, float g0, float g1,
                                            float g2, const float4 qq, float sq0, float sq1,
                                            float sq2, float rq0, float rq1, float rq2,
                                            const AxiParams &A, float alpha)
{
    if (A.on) {
        float v0, v1, v2;
        frame_axis(qq, v0, v1, v2);
        cmf_p_voxel_axi(h0, h1, h2, g0, g1, g2, v0, v1, v2, A, alpha);
        return;
    }
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
