// Shared definitions for the gfx950 HIP kernels of the tubular-segmentation
// path: status handling, the library context, launch helpers and the device
// quaternion math. Every kernel works on C-order volumes (D0, D1, D2), axis 2
// fastest, linear index i = i2 + D2 * (i1 + D1 * i0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <string>
#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <type_traits>
#include <utility>

#include "../../include/miccai_hip.h"

#define MH_EXPORT extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------- context ---

struct mh_ctx {
    int device = 0;
    std::string last_error;
    // small device block for reductions and solver control (lives as long as
    // the context): [0, 4 KiB) scalars, then partials
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // job descriptor table of mh_cmf_jobs_step (grown on demand)
    void *desc = nullptr;
    size_t desc_bytes = 0;
    // schedule of mh_cmf_solve_batched: solves per thread group and
    // iterations per visit of a group (0: chosen automatically)
    int batch_group = 0;
    int batch_visit = 0;
    int batch_sweep = 0;  // two-level sweep of batched groups: 0 automatic, 1 off, 2 on
    int gvf_sweep = 0;    // two GVF steps per launch: 0 automatic, 1 off, 2 whenever possible
    // two-level sweep of mh_cmf_solve: 0 automatic, 1 off, 2 whenever eligible
    int sweep_mode = 0;
    int sweep_shape = 0;   // tile shape (1-based; 0: default)
    int sweep_zchunk = 0;  // planes per work-group (0: modelled)
    // graph replay of the per-iteration kernels' criterion cycles: 1 on, 0 off
    int graph_mode = 0;
    hipStream_t graph_stream = nullptr;  // capture stream, created on first use
    // method the tuner chose per volume dims: (tile shape or -1 for the plain
    // kernels, plane chunk)
    std::map<std::array<int, 3>, std::pair<int, int>> sweep_tuned;
    // volume dims mh_cmf_solve has met on this context (it tunes a shape
    // the second time it meets it)
    std::set<std::array<int, 3>> sweep_seen;
    // method the batched solver's tuner chose per volume dims (tile shape,
    // plane chunk), timed on pairs of solves
    std::map<std::array<int, 3>, std::pair<int, int>> batch_tuned;
    // resident work-groups of the sweep kernel on the device, per tile shape
    // (CU count x occupancy; 0 until queried)
    std::array<int, 16> sweep_slots{};
    // this device's copy of the no-allocation solver's scratch array (set by
    // mh_create; never freed, it lives as long as the process)
    void *noalloc_scratch = nullptr;
    std::string noalloc_error;  // why it is null, if it is
};

// Resolves the no-allocation scratch array of ctx->device (current) into
// ctx->noalloc_scratch; maxflow.hip.
int mh_noalloc_setup(mh_ctx *ctx);

// Nearest marked voxel by the separable distance transform, isotropic voxel
// units (directions.hip): work holds 5 x N unsigned, [ridx | dA | bA | dB | bB];
// ridx marks voxels with their indices (0xffffffff elsewhere); dA receives the
// squared distances, bA the nearest index (lowest among equidistant ones).
int mh_nearest_marked(mh_ctx *ctx, const int64_t dims[3], unsigned *work, hipStream_t s);

namespace mh {

constexpr int kBlock = 256;  // threads per work-group for 1-D kernels (4 waves)

// Calls f(std::integral_constant<int, K>{}) for K = 0 .. N-1, in order, while
// it returns true (a loop body unrolled with its index as a compile-time
// constant); returns false if f stopped it.
template <int N, int K = 0, typename F>
__device__ __forceinline__ bool unrolled_while(F &&f)
{
    if constexpr (K < N) {
        if (!f(std::integral_constant<int, K>{})) return false;
        return unrolled_while<N, K + 1>(f);
    } else {
        return true;
    }
}

inline int set_error(mh_ctx *ctx, int code, const char *where, const char *what)
{
    if (ctx) {
        char buf[512];
        snprintf(buf, sizeof(buf), "%s: %s", where, what);
        ctx->last_error = buf;
    }
    return code;
}

inline int check_hip(mh_ctx *ctx, hipError_t e, const char *where)
{
    if (e == hipSuccess) return MH_OK;
    return set_error(ctx, MH_ERR_HIP, where, hipGetErrorString(e));
}

#define MH_CHECK(ctx, expr)                                                    \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) return mh::check_hip((ctx), e_, #expr);          \
    } while (0)

#define MH_CHECK_LAUNCH(ctx, name)                                             \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) return mh::check_hip((ctx), e_, name);           \
    } while (0)

#define MH_REQUIRE(ctx, cond, msg)                                             \
    do {                                                                       \
        if (!(cond)) return mh::set_error((ctx), MH_ERR_ARG, __func__, msg);   \
    } while (0)

inline bool dims_ok(const int64_t *dims)
{
    return dims && dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1 &&
           dims[0] * dims[1] * dims[2] < (int64_t(1) << 31);
}

__host__ __device__ inline unsigned blocks_for(int64_t n, int block = kBlock)
{
    return unsigned((n + block - 1) / block);
}

// Blocks are dealt round-robin to the 8 XCDs (each with its own L2). Remap
// the hardware block index so that XCD x works on the x-th contiguous eighth
// of the grid: stencil neighbours then mostly hit the same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned per = nb / 8u;
    if (per == 0u || b >= per * 8u) return b;
    return (b & 7u) * per + (b >> 3);
}

// -------------------------------------------------------------- quaternions --
// A frame is stored "squeezed" as z-x-z Euler angles (phi, theta, psi); the
// unit quaternion (w, x, y, z) of R = Rz(phi) Rx(theta) Rz(psi) is
//   w = cos((phi+psi)/2) cos(theta/2)   x = cos((phi-psi)/2) sin(theta/2)
//   y = sin((phi-psi)/2) sin(theta/2)   z = sin((phi+psi)/2) cos(theta/2).
// Columns of R: column 0 is the main (vessel) direction.

struct Quat {
    float w, x, y, z;
};

__device__ __forceinline__ Quat expand_squeezed(float phi, float theta, float psi)
{
    const float hp = (phi + psi) * 0.5f;
    const float hm = (phi - psi) * 0.5f;
    float st, ct, sp, cp, sm, cm;
    sincosf(theta * 0.5f, &st, &ct);
    sincosf(hp, &sp, &cp);
    sincosf(hm, &sm, &cm);
    return Quat{cp * ct, cm * st, sm * st, sp * ct};
}

// Row-major rotation matrix of a unit quaternion.
struct Mat3 {
    float m[9];
};

__device__ __forceinline__ Mat3 quat_to_mat(const Quat q)
{
#pragma clang fp contract(off)
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    Mat3 r;
    r.m[0] = 1.f - 2.f * (yy + zz); r.m[1] = 2.f * (xy - wz);       r.m[2] = 2.f * (xz + wy);
    r.m[3] = 2.f * (xy + wz);       r.m[4] = 1.f - 2.f * (xx + zz); r.m[5] = 2.f * (yz - wx);
    r.m[6] = 2.f * (xz - wy);       r.m[7] = 2.f * (yz + wx);       r.m[8] = 1.f - 2.f * (xx + yy);
    return r;
}

__device__ __forceinline__ float safe_atan2(float y, float x)
{
    return (y == 0.f && x == 0.f) ? 0.f : atan2f(y, x);
}

// Unit quaternion (w, x, y, z) -> squeezed z-x-z angles, with the two
// gimbal-lock branches (theta near 0: x^2 + y^2 < eps; theta near pi:
// 1 - (x^2 + y^2) < eps) that pin psi by convention.
__device__ __forceinline__ void squeeze_quat(float w, float x, float y, float z, float out[3])
{
    const float s = x * x + y * y;
    const bool near0 = s < FLT_EPSILON;
    const bool gimbal = near0 || (1.f - s) < FLT_EPSILON;
    const float phi = safe_atan2(x * z + y * w, x * w - y * z);
    float theta, psi;
    if (gimbal) {
        theta = near0 ? 0.f : float(M_PI);
        const float a = safe_atan2(x * y - w * z, 0.5f - (y * y + z * z));
        psi = -a + (near0 ? -phi : phi);
    } else {
        theta = acosf(fminf(fmaxf(1.f - 2.f * s, -1.f), 1.f));
        psi = safe_atan2(x * z - y * w, x * w + y * z);
    }
    out[0] = phi; out[1] = theta; out[2] = psi;
}

}  // namespace mh
