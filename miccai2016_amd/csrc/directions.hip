// Direction adjustment on gfx950: ridge detection in the vesselness cross
// section, main-direction extraction, nearest-ridge sign propagation and
// direction -> frame completion. Host-side hysteresis flood fill at the end.
#include <vector>

#include "mh_common.h"

namespace {

// A voxel is a ridge voxel when its response exceeds the threshold and the
// maxima of 8 trilinear samples taken in the plane spanned by the frame's
// second and third columns, at offsets a*v2 + b*v3 with (a, b) in
// {-1, 0, 1}^2 minus the origin. A sample's unit-cube origin is the position
// truncated toward zero; corners outside the volume contribute the value -1.
// One rounding per operation (no contraction), in the order of the
// reference's kernel, so that only the frame's trigonometry can move a
// decision, and only where the response ties a sample within rounding.
__global__ void __launch_bounds__(mh::kBlock)
ridge_kernel(float *__restrict__ ridge, const float *__restrict__ resp,
             const float *__restrict__ qeig3, int d0, int d1, int d2, float thr)
{
#pragma clang fp contract(off)
    const int64_t n = int64_t(d0) * d1 * d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int i2 = int(i % d2);
    const int i1 = int((i / d2) % d1);
    const int i0 = int(i / (int64_t(d2) * d1));
    const mh::Quat q = mh::expand_squeezed(qeig3[3 * i], qeig3[3 * i + 1], qeig3[3 * i + 2]);
    const mh::Mat3 R = mh::quat_to_mat(q);
    const float v2[3] = {R.m[1], R.m[4], R.m[7]};
    const float v3[3] = {R.m[2], R.m[5], R.m[8]};
    const float pos[3] = {float(i0), float(i1), float(i2)};
    const int dim[3] = {d0, d1, d2};
    // (a, b) offsets in the order the 8 neighbours are visited
    const int oa[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    const int ob[8] = {-1, 0, 1, -1, 1, 1, -1, 0};
    float best = -INFINITY;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float f[3];
        int o[3];
        for (int k = 0; k < 3; ++k) {
            const float p = pos[k] + float(oa[j]) * v2[k] + float(ob[j]) * v3[k];
            o[k] = int(p);  // truncation toward zero
            f[k] = p - float(o[k]);
        }
        float acc = 0.f;
        // a sample whose unit cube lies inside the volume (every sample of a
        // voxel 2 from the faces does) takes the same sums without the
        // per-corner bounds tests
        const bool inside8 = o[0] >= 0 && o[0] < d0 - 1 && o[1] >= 0 && o[1] < d1 - 1 && o[2] >= 0 &&
                             o[2] < d2 - 1;
        if (inside8) {
            const float *r0 = resp + (o[2] + int64_t(d2) * (o[1] + int64_t(d1) * o[0]));
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int cr[3] = {c >> 2, (c >> 1) & 1, c & 1};
                const float w0 = cr[0] ? f[0] : 1.f - f[0];
                const float w1 = cr[1] ? f[1] : 1.f - f[1];
                const float w2 = cr[2] ? f[2] : 1.f - f[2];
                const float s = r0[cr[2] + d2 * (cr[1] + d1 * cr[0])];
                acc += s * w0 * w1 * w2;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int cr[3] = {c >> 2, (c >> 1) & 1, c & 1};
                int ca[3];
                bool inside = true;
                float w[3];
                for (int k = 0; k < 3; ++k) {
                    ca[k] = o[k] + cr[k];
                    inside = inside && ca[k] >= 0 && ca[k] < dim[k];
                    w[k] = cr[k] ? f[k] : 1.f - f[k];
                }
                const float s =
                    inside ? resp[ca[2] + int64_t(d2) * (ca[1] + int64_t(d1) * ca[0])] : -1.f;
                acc += s * w[0] * w[1] * w[2];
            }
        }
        best = acc > best ? acc : best;
    }
    const float r = resp[i];
    ridge[i] = (r > thr && r > best) ? 1.f : 0.f;
}

__global__ void __launch_bounds__(mh::kBlock)
main_dir_kernel(float *dirs, const float *qeig3, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const mh::Quat q = mh::expand_squeezed(qeig3[3 * i], qeig3[3 * i + 1], qeig3[3 * i + 2]);
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float c0 = 1.f - 2.f * (yy + zz);
    const float c1 = 2.f * (q.x * q.y + q.w * q.z);
    const float c2 = 2.f * (q.x * q.z - q.w * q.y);
    (void)xx;
    dirs[3 * i] = c0;
    dirs[3 * i + 1] = c1;
    dirs[3 * i + 2] = c2;
}

// Brute-force nearest ridge point per voxel, ridge tiles staged in LDS. The
// scan visits ridge points in index order with a strict '<', so the lowest
// index wins ties; then the voxel's direction is flipped if it points away
// from that ridge point's direction. Distances and the product round once
// per operation, left to right (no fused multiply-adds), so a direction
// nearly perpendicular to its ridge point's is decided the same way by every
// build.
constexpr int kRidgeTile = 1024;

__global__ void __launch_bounds__(mh::kBlock)
nonridge_kernel(float *__restrict__ dirs, int d0, int d1, int d2,
                const uint16_t *__restrict__ coords, const float *__restrict__ rdirs, int nr,
                float s0, float s1, float s2)
{
#pragma clang fp contract(off)
    __shared__ uint16_t tc[kRidgeTile][3];
    const int64_t n = int64_t(d0) * d1 * d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    const bool active = i < n;
    int p0 = 0, p1 = 0, p2 = 0;
    if (active) {
        p2 = int(i % d2);
        p1 = int((i / d2) % d1);
        p0 = int(i / (int64_t(d2) * d1));
    }
    float best = INFINITY;
    int bj = 0;
    for (int t0 = 0; t0 < nr; t0 += kRidgeTile) {
        const int cnt = min(kRidgeTile, nr - t0);
        __syncthreads();
        for (int k = threadIdx.x; k < cnt * 3; k += blockDim.x) tc[k / 3][k % 3] = coords[3 * int64_t(t0) + k];
        __syncthreads();
        if (active) {
            for (int j = 0; j < cnt; ++j) {
                const float a = float(abs(p0 - int(tc[j][0]))) * s0;
                const float c = float(abs(p1 - int(tc[j][1]))) * s1;
                const float e = float(abs(p2 - int(tc[j][2]))) * s2;
                const float dsq = a * a + c * c + e * e;
                if (dsq < best) {
                    best = dsq;
                    bj = t0 + j;
                }
            }
        }
    }
    if (!active) return;
    float *v = dirs + 3 * i;
    const float *r = rdirs + 3 * int64_t(bj);
    const float x = v[0], y = v[1], z = v[2];
    if (x * r[0] + y * r[1] + z * r[2] < 0.f) {
        v[0] = -x;
        v[1] = -y;
        v[2] = -z;
    }
}

// ---- nearest ridge point by an exact separable distance transform ---------
// For an isotropic voxel size the squared distances are integers (times s^2),
// so the nearest ridge point can be found axis by axis: after the pass along
// axis a, every voxel holds the smallest squared distance to a ridge point
// reachable by moving along the axes processed so far, and the lowest ridge
// index among the points at that distance. The lowest index over a tie set
// composes across passes (each candidate already carries the lowest index of
// its own tie set), so the result equals the brute-force scan with strict '<'.
constexpr unsigned kNoRidge = 0xffffffffu;

__global__ void __launch_bounds__(mh::kBlock) fill_u32_kernel(unsigned *__restrict__ x, int64_t n, unsigned v)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) x[t] = v;
}

__global__ void __launch_bounds__(mh::kBlock)
ridge_scatter_kernel(unsigned *__restrict__ ridx, const uint16_t *__restrict__ coords, int nr,
                     int d0, int d1, int d2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    const int c0 = coords[3 * j], c1 = coords[3 * j + 1], c2 = coords[3 * j + 2];
    if (c0 >= d0 || c1 >= d1 || c2 >= d2) return;
    atomicMin(&ridx[c2 + int64_t(d2) * (c1 + int64_t(d1) * c0)], unsigned(j));
}

// Pass along axis 2: two sweeps per line give the nearest ridge voxel on the
// line (the lower index wins when the left and right ones are equidistant).
__global__ void __launch_bounds__(mh::kBlock)
edt_axis2_kernel(const unsigned *__restrict__ ridx, unsigned *__restrict__ dist,
                 unsigned *__restrict__ best, int d0, int d1, int d2)
{
    const int64_t line = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (line >= int64_t(d0) * d1) return;
    const int64_t base = line * d2;
    int last = -1;
    unsigned lastj = kNoRidge;
    for (int x = 0; x < d2; ++x) {
        const unsigned j = ridx[base + x];
        if (j != kNoRidge) { last = x; lastj = j; }
        if (last >= 0) {
            const unsigned dx = unsigned(x - last);
            dist[base + x] = dx * dx;
            best[base + x] = lastj;
        } else {
            dist[base + x] = kNoRidge;
            best[base + x] = kNoRidge;
        }
    }
    last = -1;
    for (int x = d2 - 1; x >= 0; --x) {
        const unsigned j = ridx[base + x];
        if (j != kNoRidge) { last = x; lastj = j; }
        if (last >= 0) {
            const unsigned dx = unsigned(last - x);
            const unsigned dd = dx * dx;
            const unsigned cur = dist[base + x];
            if (dd < cur || (dd == cur && lastj < best[base + x])) {
                dist[base + x] = dd;
                best[base + x] = lastj;
            }
        }
    }
}

// The same pass with one wave per line, the line in registers (E elements
// per lane, lane l holding positions l*E .. l*E+E-1): the nearest marked
// voxel at or left of each position is a prefix maximum of marked positions,
// the nearest at or right of it a suffix minimum, both over the wave by
// shuffles; the lower index wins when the left and right ones are
// equidistant.
template <int E>
__global__ void __launch_bounds__(256)
edt_axis2_wave_kernel(const unsigned *__restrict__ ridx, unsigned *__restrict__ dist,
                      unsigned *__restrict__ best, int64_t nlines, int d2)
{
    const int64_t line = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (line >= nlines) return;  // a whole wave: the line is the wave's
    const int lane = int(threadIdx.x & 63);
    const unsigned *in = ridx + line * d2;
    int pre[E], suf[E];
    int run = -1;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int x = lane * E + k;
        const bool m = x < d2 && in[x] != kNoRidge;
        run = m ? x : run;
        pre[k] = run;
    }
    // exclusive prefix maximum of the lanes' last marked positions
    int inc = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off, 64);
        if (lane >= off) inc = max(inc, v);
    }
    int exl = __shfl_up(inc, 1, 64);
    if (lane == 0) exl = -1;
    run = 0x7fffffff;
#pragma unroll
    for (int k = E - 1; k >= 0; --k) {
        const int x = lane * E + k;
        const bool m = x < d2 && in[x] != kNoRidge;
        run = m ? x : run;
        suf[k] = run;
        pre[k] = max(pre[k], exl);
    }
    // exclusive suffix minimum of the lanes' first marked positions
    int sinc = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_down(sinc, off, 64);
        if (lane + off < 64) sinc = min(sinc, v);
    }
    int sexl = __shfl_down(sinc, 1, 64);
    if (lane == 63) sexl = 0x7fffffff;
    unsigned *od = dist + line * d2, *ob = best + line * d2;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int x = lane * E + k;
        if (x >= d2) break;
        const int f = pre[k], b = min(suf[k], sexl);
        int c = f;
        if (b < 0x7fffffff) {
            if (f < 0) {
                c = b;
            } else {
                const unsigned df = unsigned(x - f) * unsigned(x - f);
                const unsigned db = unsigned(b - x) * unsigned(b - x);
                if (db < df || (db == df && in[b] < in[f])) c = b;
            }
        }
        if (c < 0) {
            od[x] = kNoRidge;
            ob[x] = kNoRidge;
        } else {
            const unsigned dx = unsigned(abs(x - c));
            od[x] = dx * dx;
            ob[x] = in[c];
        }
    }
}

// Pass along axis 1 or 0: scan outward from the voxel, stopping once the
// squared offset alone reaches the best value found.
template <int AXIS>
__global__ void __launch_bounds__(mh::kBlock)
edt_scan_kernel(const unsigned *__restrict__ din, const unsigned *__restrict__ bin,
                unsigned *__restrict__ dout, unsigned *__restrict__ bout, int d0, int d1, int d2)
{
    const int64_t n = int64_t(d0) * d1 * d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int len = AXIS == 1 ? d1 : d0;
    const int64_t st = AXIS == 1 ? int64_t(d2) : int64_t(d1) * d2;
    const int c = AXIS == 1 ? int((i / d2) % d1) : int(i / (int64_t(d2) * d1));
    unsigned bd = din[i], bj = bin[i];
    for (int off = 1; off < len; ++off) {
        const unsigned o2 = unsigned(off) * unsigned(off);
        if (o2 > bd) break;  // every further candidate is at least o2 away
        if (c - off >= 0) {
            const unsigned v = din[i - off * st];
            if (v != kNoRidge) {
                const unsigned dd = v + o2;
                const unsigned j = bin[i - off * st];
                if (dd < bd || (dd == bd && j < bj)) { bd = dd; bj = j; }
            }
        }
        if (c + off < len) {
            const unsigned v = din[i + off * st];
            if (v != kNoRidge) {
                const unsigned dd = v + o2;
                const unsigned j = bin[i + off * st];
                if (dd < bd || (dd == bd && j < bj)) { bd = dd; bj = j; }
            }
        }
    }
    dout[i] = bd;
    bout[i] = bj;
}

__global__ void __launch_bounds__(mh::kBlock)
flip_by_nearest_kernel(float *__restrict__ dirs, const unsigned *__restrict__ best,
                       const float *__restrict__ rdirs, int64_t n)
{
#pragma clang fp contract(off)
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned j = best[i];
    if (j == kNoRidge) return;
    float *v = dirs + 3 * i;
    const float *r = rdirs + 3 * int64_t(j);
    const float x = v[0], y = v[1], z = v[2];
    if (x * r[0] + y * r[1] + z * r[2] < 0.f) {
        v[0] = -x;
        v[1] = -y;
        v[2] = -z;
    }
}

__device__ __forceinline__ void normalize3(float v[3])
{
    const float d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (d == 0.f) return;  // a zero vector stays zero
    const float r = 1.f / sqrtf(d);
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
}

// Quaternion (w, x, y, z) of a rotation matrix by the largest-diagonal
// branch (trace > 0, else the largest of the three diagonal entries).
__device__ __forceinline__ mh::Quat mat_to_quat(const float m[9])
{
    mh::Quat q;
    const float tr = m[0] + m[4] + m[8];
    if (tr > 0.f) {
        const float s = 0.5f * rsqrtf(tr + 1.f);
        q.w = 0.25f / s;
        q.x = (m[7] - m[5]) * s;
        q.y = (m[2] - m[6]) * s;
        q.z = (m[3] - m[1]) * s;
    } else if (m[0] > m[4] && m[0] > m[8]) {
        const float s = 0.5f * rsqrtf(1.f + m[0] - m[4] - m[8]);
        q.w = (m[7] - m[5]) * s;
        q.x = 0.25f / s;
        q.y = (m[1] + m[3]) * s;
        q.z = (m[2] + m[6]) * s;
    } else if (m[4] > m[8]) {
        const float s = 0.5f * rsqrtf(1.f + m[4] - m[0] - m[8]);
        q.w = (m[2] - m[6]) * s;
        q.x = (m[1] + m[3]) * s;
        q.y = 0.25f / s;
        q.z = (m[5] + m[7]) * s;
    } else {
        const float s = 0.5f * rsqrtf(1.f + m[8] - m[0] - m[4]);
        q.w = (m[3] - m[1]) * s;
        q.x = (m[2] + m[6]) * s;
        q.y = (m[5] + m[7]) * s;
        q.z = 0.25f / s;
    }
    return q;
}

// v -> (e0, e1, e0 x e1) with e0 = v/|v| and e1 perpendicular to e0 in the
// coordinate plane that excludes e0's larger of |x|, |y|; det = +1.
__global__ void __launch_bounds__(mh::kBlock)
frame_kernel(float *out3, const float *dirs3, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float e0[3] = {dirs3[3 * i], dirs3[3 * i + 1], dirs3[3 * i + 2]};
    normalize3(e0);
    float e1[3];
    if (fabsf(e0[0]) > fabsf(e0[1])) {
        e1[0] = -e0[2]; e1[1] = 0.f; e1[2] = e0[0];
    } else {
        e1[0] = 0.f; e1[1] = e0[2]; e1[2] = -e0[1];
    }
    normalize3(e1);
    const float e2[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                         e0[0] * e1[1] - e0[1] * e1[0]};
    const float m[9] = {e0[0], e1[0], e2[0], e0[1], e1[1], e2[1], e0[2], e1[2], e2[2]};
    const mh::Quat q = mat_to_quat(m);
    float s[3];
    mh::squeeze_quat(q.w, q.x, q.y, q.z, s);
    out3[3 * i] = s[0];
    out3[3 * i + 1] = s[1];
    out3[3 * i + 2] = s[2];
}

}  // namespace

MH_EXPORT int mh_detect_ridges(mh_ctx *ctx, float *ridge, const float *resp, const float *qeig3,
                               const int64_t dims[3], float thr, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, ridge && resp && qeig3 && ridge != resp, "bad pointers");
    const int64_t n = dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(ridge_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, ridge, resp, qeig3, int(dims[0]), int(dims[1]),
                       int(dims[2]), thr);
    MH_CHECK_LAUNCH(ctx, "ridge_kernel");
    return MH_OK;
}

MH_EXPORT int mh_main_dirs(mh_ctx *ctx, float *dirs, const float *qeig3, int64_t n, void *stream)
{
    MH_REQUIRE(ctx, dirs && qeig3 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(main_dir_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, dirs, qeig3, n);
    MH_CHECK_LAUNCH(ctx, "main_dir_kernel");
    return MH_OK;
}

MH_EXPORT int64_t mh_nonridge_work_bytes(const int64_t dims[3])
{
    if (!mh::dims_ok(dims)) return -1;
    return int64_t(5) * dims[0] * dims[1] * dims[2] * int64_t(sizeof(unsigned));
}

MH_EXPORT int mh_nonridge_signs(mh_ctx *ctx, float *dirs, const int64_t dims[3],
                                const uint16_t *coords, const float *rdirs, int64_t nr,
                                const float scaling[3], void *work, int64_t work_bytes,
                                void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, dirs && scaling && nr >= 0 && nr < (int64_t(1) << 31), "bad arguments");
    MH_REQUIRE(ctx, nr == 0 || (coords && rdirs), "null ridge arrays");
    if (nr == 0) return MH_OK;  // no ridge point: nothing to align with
    const int64_t n = dims[0] * dims[1] * dims[2];
    hipStream_t s = (hipStream_t)stream;
    const bool iso = scaling[0] == scaling[1] && scaling[1] == scaling[2] && scaling[0] > 0.f;
    const bool small_dims = dims[0] <= 16384 && dims[1] <= 16384 && dims[2] <= 16384;
    const bool have_ws = work && work_bytes >= mh_nonridge_work_bytes(dims);
    // the brute-force scan costs N*R; the transform about N*(nearest distance)
    if (!(iso && small_dims && have_ws) || double(n) * double(nr) < 2e9) {
        hipLaunchKernelGGL(nonridge_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0, s, dirs,
                           int(dims[0]), int(dims[1]), int(dims[2]), coords, rdirs, int(nr),
                           scaling[0], scaling[1], scaling[2]);
        MH_CHECK_LAUNCH(ctx, "nonridge_kernel");
        return MH_OK;
    }
    unsigned *ridx = static_cast<unsigned *>(work);
    unsigned *dA = ridx + n, *bA = dA + n;
    const unsigned nb = mh::blocks_for(n);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(nb), dim3(mh::kBlock), 0, s, ridx, n, kNoRidge);
    hipLaunchKernelGGL(ridge_scatter_kernel, dim3(mh::blocks_for(nr)), dim3(mh::kBlock), 0, s, ridx,
                       coords, int(nr), int(dims[0]), int(dims[1]), int(dims[2]));
    MH_CHECK_LAUNCH(ctx, "ridge_scatter_kernel");
    const int rc = mh_nearest_marked(ctx, dims, static_cast<unsigned *>(work), s);
    if (rc != MH_OK) return rc;
    hipLaunchKernelGGL(flip_by_nearest_kernel, dim3(nb), dim3(mh::kBlock), 0, s, dirs, bA, rdirs, n);
    MH_CHECK_LAUNCH(ctx, "nearest-ridge transform");
    return MH_OK;
}

// The separable transform over work = [ridx | dA | bA | dB | bB] (N unsigned
// each): ridx holds every marked voxel's index (kNoRidge elsewhere); on
// return dA holds each voxel's squared distance (voxel units) to the nearest
// marked voxel and bA that voxel's index, the lowest among equidistant ones
// (kNoRidge where nothing is marked).
int mh_nearest_marked(mh_ctx *ctx, const int64_t dims[3], unsigned *work, hipStream_t s)
{
    const int64_t n = dims[0] * dims[1] * dims[2];
    unsigned *ridx = work;
    unsigned *dA = ridx + n, *bA = dA + n, *dB = bA + n, *bB = dB + n;
    const int d0 = int(dims[0]), d1 = int(dims[1]), d2 = int(dims[2]);
    const unsigned nb = mh::blocks_for(n);
    const int64_t nlines = int64_t(d0) * d1;
    const unsigned nbw = unsigned((nlines * 64 + 255) / 256);
    if (d2 <= 256) {
        hipLaunchKernelGGL(edt_axis2_wave_kernel<4>, dim3(nbw), dim3(256), 0, s, ridx, dA, bA, nlines, d2);
    } else if (d2 <= 512) {
        hipLaunchKernelGGL(edt_axis2_wave_kernel<8>, dim3(nbw), dim3(256), 0, s, ridx, dA, bA, nlines, d2);
    } else if (d2 <= 1024) {
        hipLaunchKernelGGL(edt_axis2_wave_kernel<16>, dim3(nbw), dim3(256), 0, s, ridx, dA, bA, nlines, d2);
    } else {
        hipLaunchKernelGGL(edt_axis2_kernel, dim3(mh::blocks_for(int64_t(d0) * d1)), dim3(mh::kBlock), 0,
                           s, ridx, dA, bA, d0, d1, d2);
    }
    hipLaunchKernelGGL(edt_scan_kernel<1>, dim3(nb), dim3(mh::kBlock), 0, s, dA, bA, dB, bB, d0, d1, d2);
    hipLaunchKernelGGL(edt_scan_kernel<0>, dim3(nb), dim3(mh::kBlock), 0, s, dB, bB, dA, bA, d0, d1, d2);
    MH_CHECK_LAUNCH(ctx, "separable nearest-voxel transform");
    return MH_OK;
}

MH_EXPORT int mh_complete_frames(mh_ctx *ctx, float *out3, const float *dirs3, int64_t n,
                                 void *stream)
{
    MH_REQUIRE(ctx, out3 && dirs3 && n >= 0 && n < (int64_t(1) << 31), "bad arguments");
    if (n == 0) return MH_OK;
    hipLaunchKernelGGL(frame_kernel, dim3(mh::blocks_for(n)), dim3(mh::kBlock), 0,
                       (hipStream_t)stream, out3, dirs3, n);
    MH_CHECK_LAUNCH(ctx, "frame_kernel");
    return MH_OK;
}

// Breadth-first flood fill over the 26-neighbourhood, restricted to
// candidate voxels, from every seed (seed = sure && cand).
MH_EXPORT int mh_hysteresis_host(const uint8_t *cand, const uint8_t *sure, uint8_t *out,
                                 const int64_t dims[3])
{
    if (!cand || !sure || !out || !mh::dims_ok(dims)) return MH_ERR_ARG;
    const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2];
    const int64_t n = d0 * d1 * d2;
    std::vector<int64_t> queue;
    for (int64_t i = 0; i < n; ++i) {
        out[i] = 0;
    }
    for (int64_t i = 0; i < n; ++i)
        if (sure[i] && cand[i]) {
            out[i] = 1;
            queue.push_back(i);
        }
    for (size_t h = 0; h < queue.size(); ++h) {
        const int64_t i = queue[h];
        const int64_t i2 = i % d2, i1 = (i / d2) % d1, i0 = i / (d1 * d2);
        for (int64_t a = std::max<int64_t>(0, i0 - 1); a <= std::min(d0 - 1, i0 + 1); ++a)
            for (int64_t bb = std::max<int64_t>(0, i1 - 1); bb <= std::min(d1 - 1, i1 + 1); ++bb)
                for (int64_t c = std::max<int64_t>(0, i2 - 1); c <= std::min(d2 - 1, i2 + 1); ++c) {
                    const int64_t j = c + d2 * (bb + d1 * a);
                    if (cand[j] && !out[j]) {
                        out[j] = 1;
                        queue.push_back(j);
                    }
                }
    }
    return MH_OK;
}
