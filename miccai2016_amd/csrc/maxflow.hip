// Anisotropic-TV continuous max-flow on gfx950.
//
// State per voxel: hp = H p (3 floats, H = R(q) diag(sqrt(lambda))), the sink
// flow pt, the labelling u, and tmp = div(hp) - ps + pt - rcc u, the residual
// the p-step differentiates. One iteration is two stencil launches:
//   p-step: p = H^-1 hp + H^T (ss * grad+ tmp); project |p| <= alpha; hp = H p
//   u-step: div hp; ps, pt projections; err; u -= err / rcc; tmp = err - rcc u
// The reference's third launch (tmp -= rcc u) is folded into the u-step; on
// criterion iterations the u-step also stores per-block sums of |err| and a
// one-block kernel adds them in block order and decides the stop, so the host
// never waits inside the loop. A stopped solve turns every later launch into
// an immediate return.
#include <mutex>
#include <vector>

#include "cmf_voxel.h"
#include "mh_common.h"
#include "mh_reduce.h"

namespace {

// control block (int32 words) at the head of the solver workspace
enum : int { kStop = 0, kNcrit = 1, kCounter = 2, kParity = 3, kCtrlWords = 64 };
constexpr int kCritLogMax = 1 << 16;

struct SumOp {
    __device__ static float identity() { return 0.f; }
    __device__ static float combine(float a, float b) { return a + b; }
};

struct Dims {
    int d0, d1, d2;
};


__device__ __forceinline__ void coords(int64_t i, const Dims D, int &i0, int &i1, int &i2)
{
    i2 = int(i % D.d2);
    const int64_t r = i / D.d2;
    i1 = int(r % D.d1);
    i0 = int(r / D.d1);
}

// ------------------------------------------------------------------ init ---

// q4 = unit quaternion of the squeezed frame; pt = min(cs, ct); hp = 0;
// tmp = -rcc u (the residual with ps = pt and div hp = 0).
__global__ void __launch_bounds__(mh::kBlock)
cmf_init_kernel(float4 *__restrict__ q4, const float *__restrict__ eigs, int eig_stride,
                int eig_offset, float *__restrict__ hp, float *__restrict__ pt,
                float *__restrict__ tmp, const float *__restrict__ u, const float *__restrict__ cs,
                const float *__restrict__ ct, int64_t n, float rcc)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *e = eigs + i * eig_stride + eig_offset;
    const mh::Quat q = mh::expand_squeezed(e[0], e[1], e[2]);
    q4[i] = make_float4(q.w, q.x, q.y, q.z);
    hp[i] = 0.f;
    hp[n + i] = 0.f;
    hp[2 * n + i] = 0.f;
    pt[i] = fminf(cs[i], ct[i]);
    tmp[i] = -(u[i] * rcc);
}

// ---------------------------------------------------------------- p-step ---

struct PParams {
    float ss0, ss1, ss2;
    float sq0, sq1, sq2;  // constant sqrt(lambda) (CONST_EV)
    float rq0, rq1, rq2;  // 1 / sq (CONST_EV)
    mh::AxiParams A;      // the axial p-step (CONST_EV, l1 == l2)
};

PParams p_params(const mh_cmf_params *prm)
{
    return PParams{prm->ss[0], prm->ss[1], prm->ss[2], prm->sqev[0], prm->sqev[1], prm->sqev[2],
                   float(1.0 / double(prm->sqev[0])), float(1.0 / double(prm->sqev[1])),
                   float(1.0 / double(prm->sqev[2])), mh::axi_params(prm)};
}

template <bool CONST_EV>
__global__ void __launch_bounds__(mh::kBlock)
cmf_p_kernel(float *__restrict__ hp, const float *__restrict__ tmp, const float4 *__restrict__ q4,
             const float *__restrict__ eigs6, const float *__restrict__ alpha, const Dims D,
             const PParams P, const int *__restrict__ ctrl)
{
    if (ctrl[kStop]) return;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int i0, i1, i2;
    coords(i, D, i0, i1, i2);
    const int64_t s0 = int64_t(D.d1) * D.d2;
    const float t = tmp[i];
    const float g0 = i0 < D.d0 - 1 ? (tmp[i + s0] - t) * P.ss0 : 0.f;
    const float g1 = i1 < D.d1 - 1 ? (tmp[i + D.d2] - t) * P.ss1 : 0.f;
    const float g2 = i2 < D.d2 - 1 ? (tmp[i + 1] - t) * P.ss2 : 0.f;
    float sq0 = P.sq0, sq1 = P.sq1, sq2 = P.sq2, rq0 = P.rq0, rq1 = P.rq1, rq2 = P.rq2;
    if (!CONST_EV) {
        sq0 = eigs6[6 * i];
        sq1 = eigs6[6 * i + 1];
        sq2 = eigs6[6 * i + 2];
        rq0 = 1.f / sq0;
        rq1 = 1.f / sq1;
        rq2 = 1.f / sq2;
    }
    float h0 = hp[i], h1 = hp[n + i], h2 = hp[2 * n + i];
    mh::AxiParams A = P.A;
    if (!CONST_EV) A.on = 0;  // pointwise eigenvalues: the general form
    mh::cmf_p_voxel(h0, h1, h2, g0, g1, g2, q4[i], sq0, sq1, sq2, rq0, rq1, rq2, A, alpha[i]);
    hp[i] = h0;
    hp[n + i] = h1;
    hp[2 * n + i] = h2;
}

// ---------------------------------------------------------------- u-step ---

struct UParams {
    float rcc, inv_rcc;
    float rs0, rs1, rs2;
    float error_bound;
    float inv_n_rcc;  // 1 / (N rcc)
};

template <bool SUP, bool CHECK>
__global__ void __launch_bounds__(mh::kBlock)
cmf_u_kernel(const float *__restrict__ hp, float *__restrict__ tmp, const float *__restrict__ cs,
             const float *__restrict__ ct, float *__restrict__ pt, float *__restrict__ u,
             const uint8_t *__restrict__ labels, const Dims D, const UParams P,
             int *__restrict__ ctrl, float *__restrict__ partials, float *__restrict__ crit_log)
{
    if (ctrl[kStop]) return;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    float aerr = 0.f;
    if (i < n) {
        int i0, i1, i2;
        coords(i, D, i0, i1, i2);
        const int64_t s0 = int64_t(D.d1) * D.d2;
        const float *hp0 = hp, *hp1 = hp + n, *hp2 = hp + 2 * n;
        // negative adjoint of the forward difference
        const float x0 = (i0 < D.d0 - 1 ? hp0[i] : 0.f) - (i0 > 0 ? hp0[i - s0] : 0.f);
        const float x1 = (i1 < D.d1 - 1 ? hp1[i] : 0.f) - (i1 > 0 ? hp1[i - D.d2] : 0.f);
        const float x2 = (i2 < D.d2 - 1 ? hp2[i] : 0.f) - (i2 > 0 ? hp2[i - 1] : 0.f);
        const float div = mh::cmf_div(x0, x1, x2, P.rs0, P.rs1, P.rs2);
        float ub = 1.f, uf = 0.f;
        if (SUP) {
            const uint8_t l = labels[i];
            uf = float(l & 1);
            ub = (l & 2) ? 0.f : 1.f;
        }
        float ui = u[i], pti = pt[i], ti;
        const float err = mh::cmf_u_voxel(div, pti, ui, ti, cs[i], ct[i], P.rcc, P.inv_rcc, ub, uf);
        pt[i] = pti;
        u[i] = ui;
        tmp[i] = ti;
        aerr = fabsf(err);
    }
    if (!CHECK) return;
    aerr = mh::block_reduce<SumOp>(aerr);
    if (threadIdx.x == 0) partials[blockIdx.x] = aerr;
}

// Criterion of one solve from the u-step's per-block partials (summed in
// block order): log it and set the stop flag on NaN or convergence. One
// work-group; launched after a criterion iteration's u-step.
__global__ void __launch_bounds__(1024)
cmf_crit_kernel(const float *__restrict__ partials, int nparts, float inv_n_rcc,
                float error_bound, int *__restrict__ ctrl, float *__restrict__ crit_log)
{
    if (ctrl[kStop]) return;
    const float s = mh::block_sum_array<SumOp>(partials, nparts);
    if (threadIdx.x == 0) {
        const float crit = s * inv_n_rcc;
        const int k = ctrl[kNcrit];
        if (k < kCritLogMax) crit_log[k] = crit;
        ctrl[kNcrit] = k + 1;
        if (crit != crit) ctrl[kStop] = 1;
        else if (crit < error_bound) ctrl[kStop] = 2;
    }
}

constexpr int kSI = 8;  // most solves per group of the batched solver

// an event destroyed when it goes out of scope, error returns included
struct ScopedEvent {
    hipEvent_t e = nullptr;
    ScopedEvent() = default;
    ScopedEvent(const ScopedEvent &) = delete;
    ScopedEvent &operator=(const ScopedEvent &) = delete;
    ~ScopedEvent()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// per-solve constants of a group: alpha = w exp(-|grad I|^2 k), k = 1 / decay^2
struct SolveParams {
    float w[kSI];
    float k[kSI];  // 1 / decay^2
};

// ----------------------------------------------------- two-level sweep ---
// Two iterations (levels A and B) in one pass over the state, temporally
// blocked along axis 0. A work-group owns an output tile O of TY x TX voxels
// in the (axis 1, axis 2) plane and a chunk of planes [z0, z1), and sweeps z
// upward. At step z it computes
//   1. the p-step of level A at plane z on O widened by (2, 1) voxels below
//      and above in both in-plane axes (region E3),
//   2. the u-step of level A at plane z on E3 minus its first row/column (E2),
//   3. the p-step of level B at plane z-1 on O widened by one below (E1),
//   4. the u-step of level B at plane z-1 on O, and stores the result,
// so the in-plane neighbours each step needs (p: +1, u: -1) come from the
// widened regions through LDS, and the axis-0 neighbours (plane z+1 for p,
// plane z-1 for u) from registers carried between steps. A thread owns the
// same in-plane points at every step. Level A reads the state of iteration t
// from one buffer set and level B writes iteration t+2 to the other, so
// work-groups never see each other's outputs within a pass. With levels = 1
// only steps 1 and 2 run and step 2 stores. The per-voxel arithmetic is the
// shared helpers', so the result equals two launches of the plain kernels bit
// for bit.
// Output tiles (axis 1 x axis 2), work-group sizes and levels: the widened
// tile (TY + 2L - 1) x (TX + 2L - 1) of an L-level sweep has one point per
// thread. Passes shorter than a shape's levels (a criterion or the end cuts
// them) run shape 0 (one or two levels) or shape 5 (three).
struct SweepShape {
    int ty, tx, threads, levels;
};
constexpr SweepShape kSweepShapes[] = {{12, 64, 1024, 2}, {11, 32, 512, 2}, {4, 64, 512, 2},
                                       {26, 32, 1024, 2}, {13, 61, 1024, 2}, {11, 59, 1024, 3},
                                       {27, 27, 1024, 3}, {9, 57, 1024, 4},  {25, 25, 1024, 4}};
constexpr int kNumSweepShapes = 9;
constexpr int kShape3 = 5;  // the three-level shape of a four-level method's three-level passes
constexpr int kDefaultShape = 6;  // untuned solves: the 27 x 27 three-level shape
// iterations timed per candidate method (after a warm-up pass): shorter
// windows let launch-to-launch noise of a few per cent pick a slower shape
constexpr int kTuneIters = 24;

// Two solves per thread: the descriptors span both solves' fields (hp: 24 B
// per voxel), so the volume must stay below half the single-solve limit
// (every byte offset below 2^31).
constexpr int64_t kSweepPairMaxVoxels = int64_t(1) << 26;

// Batched solves advance two per thread with the 1024-thread tile shapes
// (measured 5 % faster at 128^3: one work-group per CU either way, the pair
// sharing frames, capacities and barriers); the 512-thread shapes would drop
// from three resident work-groups per CU to two, and stay one per thread.
bool sweep_pairs(int shape_ix, int64_t n)
{
    // (four levels for two solves would need more than 128 registers)
    return kSweepShapes[shape_ix].threads >= 1024 && kSweepShapes[shape_ix].levels <= 3 &&
           n < kSweepPairMaxVoxels;
}

// Levels of the sweep pass at iteration it of a shape of `levels` levels,
// the last iteration `lim` - 1 and a criterion every `update` iterations: a
// criterion iteration can only be a pass's last level (its u-step sums the
// error), so passes are cut short before one, and at the end.
int sweep_levels(int levels, int it, int lim, int update)
{
    int lv = 1;
    while (lv < levels && it + lv < lim && (it + lv - 1) % update != 0) ++lv;
    return lv;
}

// a way to run iterations: shape < 0 the plain kernels, else a multi-level
// sweep with that tile shape and plane chunk
struct SweepMethod {
    int shape, zchunk;
};

struct SweepBufs {
    const float *hp, *tmp, *pt, *u;  // iteration t
    float *hpo, *tmpo, *pto, *uo;    // iteration t + levels
};

// A batched pass: `ns` solves of one volume in one launch, block b serving
// solve b % ns; solve k's state and alpha are at offset k n (hp: 3 k n) of
// their buffers, and it has its own stop flag, parity and criterion partials
// (nper per solve).
struct SweepBatch {
    int s0, ns, nper;
    const int *stop;
    int *parity;
};

// Inputs of one point at one plane, loaded a step ahead of their use: those
// of one solve, and those the solves of one volume share.
struct SweepIn {
    float t, t0, t1, t2;  // tmp at the voxel and its +axis-0/1/2 neighbours
    float h0, h1, h2;     // hp
    float a, pt, u;
};
struct SweepInShared {
    float4 q;
    float cs, ct;
};

// The sweep addresses its fields through buffer resources: a uniform byte
// offset (the plane, and the component of hp) plus the thread's 32-bit byte
// offset within the plane, so no 64-bit address is formed per lane. Every
// byte offset stays below 2^31 (the sweep takes volumes of fewer than 2^27
// voxels: the largest field, the AoS x4 frame, holds 16 B per voxel), so the
// offsets are non-negative as the signed operands of the buffer builtins, and
// each resource spans exactly its field. Larger volumes run the per-iteration
// kernels, which index in 64 bits.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int64_t kSweepMaxVoxels = int64_t(1) << 27;
constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit data, gfx9 family

__device__ __forceinline__ Rsrc rsrc_of(const void *p, int64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(unsigned(bytes)), kRsrcWord3);
}
__device__ __forceinline__ float bld(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}
__device__ __forceinline__ float4 bld4(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), int(soff), 0));
}
__device__ __forceinline__ void bst(Rsrc r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, int(voff), int(soff), 0);
}
// the same store through a 64-bit base address (two scalar registers instead
// of a four-register descriptor). The byte offset voff + soff is summed in 32
// bits (it stays below 2^31 for every volume the sweep takes, as the
// descriptors' offsets do), so the store takes the base in scalar registers
// and one 32-bit offset, which stores at the same offset into different
// fields share; adding soff to a base + voff address instead kept a 64-bit
// address per field in vector registers and a 64-bit add per store.
__device__ __forceinline__ void pst(float *base, unsigned voff, unsigned soff, float v)
{
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + size_t(voff + soff)) = v;
}

struct SweepRsrc {
    Rsrc tmp, hp, q, a, pt, u, cs, ct;  // inputs
    Rsrc hpo, tmpo, pto, uo;            // outputs
    // the inputs as addresses, for loads through 64-bit bases
    const float *ptmp, *php, *pa, *ppt, *pu, *pcs, *pct;
    const float4 *pq;
};

// Stores of the sweep. The 1024-thread shapes hold one work-group per CU at
// any register count up to 128, so they store through addresses: the
// descriptors they save would otherwise be spilled scalar registers. The
// 512-thread shapes keep the descriptors, which need fewer vector registers.
template <bool PTR>
__device__ __forceinline__ void sweep_store(Rsrc r, float *base, unsigned voff, unsigned soff,
                                            float v)
{
    if constexpr (PTR) pst(base, voff, soff, v);
    else bst(r, voff, soff, v);
}

// Loads of the sweep: through a descriptor (plane offset in a scalar
// register, for free), or through a 64-bit base with the 32-bit byte offset
// voff + soff (which stays below 2^31, as the descriptors' offsets do). One
// solve per thread with 1024 threads loads through bases: the descriptors'
// 32 scalar registers were spilled and reloaded in every plane step (84 -> 74
// VGPRs, 37 -> 21 lane reloads per two plane steps); with two solves per
// thread the bases measured more instructions, so the pair kernel keeps the
// descriptors.
template <bool PTR>
__device__ __forceinline__ float sweep_ld(Rsrc r, const float *p, unsigned voff, unsigned soff)
{
    if constexpr (PTR)
        return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + size_t(voff + soff));
    else return bld(r, voff, soff);
}
template <bool PTR>
__device__ __forceinline__ float4 sweep_ld4(Rsrc r, const float4 *p, unsigned voff, unsigned soff)
{
    if constexpr (PTR)
        return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(p) + size_t(voff + soff));
    else return bld4(r, voff, soff);
}

// The inputs of NS solves (solve j's scalar fields at byte offset j nb within
// the buffers, its hp at 3 j nb) and those they share, at plane zb. Every lane
// loads, with offsets kept inside the volume: the forward neighbours' offsets
// fall back to the point itself on the far faces (ob1, ob2 for axes 1 and 2,
// e0 false for axis 0), and the caller clamps the plane. Lanes and faces
// whose values are not needed load in-range values that nothing reads (the
// phases mask the same faces and regions); loading them all spares the
// branches around conditional loads and the registers that held a value for
// lanes that skipped them (one solve per thread: 708 -> 646 VALU and 548 ->
// 442 SALU instructions per kernel, 1-4 % faster by tile shape at 128^3 and
// 256^3). Two solves per thread load as sweep_load does: the extra loads of
// the halo lanes for both solves made the pair kernel 3 % slower.
template <int NS, bool PTR>
__device__ __forceinline__ void sweep_load_all(SweepIn (&L)[NS], SweepInShared &C, const SweepRsrc &R,
                                               unsigned nb, unsigned pb, unsigned zb, unsigned ob,
                                               unsigned ob1, unsigned ob2, bool e0)
{
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const unsigned zs = zb + unsigned(j) * nb, zh = zb + 3u * unsigned(j) * nb;
        L[j].t = sweep_ld<PTR>(R.tmp, R.ptmp, ob, zs);
        L[j].t0 = sweep_ld<PTR>(R.tmp, R.ptmp, ob, zs + (e0 ? pb : 0u));
        L[j].t1 = sweep_ld<PTR>(R.tmp, R.ptmp, ob1, zs);
        L[j].t2 = sweep_ld<PTR>(R.tmp, R.ptmp, ob2, zs);
        L[j].h0 = sweep_ld<PTR>(R.hp, R.php, ob, zh);
        L[j].h1 = sweep_ld<PTR>(R.hp, R.php, ob, zh + nb);
        L[j].h2 = sweep_ld<PTR>(R.hp, R.php, ob, zh + 2u * nb);
    }
    C.q = sweep_ld4<PTR>(R.q, R.pq, 4u * ob, 4u * zb);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        L[j].a = sweep_ld<PTR>(R.a, R.pa, ob, zb + unsigned(j) * nb);
        L[j].pt = sweep_ld<PTR>(R.pt, R.ppt, ob, zb + unsigned(j) * nb);
        L[j].u = sweep_ld<PTR>(R.u, R.pu, ob, zb + unsigned(j) * nb);
    }
    C.cs = sweep_ld<PTR>(R.cs, R.pcs, ob, zb);
    C.ct = sweep_ld<PTR>(R.ct, R.pct, ob, zb);
}

// The inputs of NS solves, in one branch: only the lanes of the level-A
// p-step's region (the tile's points inside the volume, at a plane inside it)
// load, the u-step's inputs too (its region lies inside; one branch instead
// of two measured 0.4 % faster for the pair kernel). The forward neighbours'
// offsets fall back to the point itself on the far faces (ob1, ob2; e0
// false), as in sweep_load_all, which the phases mask.
template <int NS>
__device__ __forceinline__ void sweep_load(SweepIn (&L)[NS], SweepInShared &C, const SweepRsrc &R,
                                           unsigned nb, unsigned pb, unsigned zb, unsigned ob,
                                           unsigned ob1, unsigned ob2, bool p_on, bool e0)
{
    if (p_on) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const unsigned zs = zb + unsigned(j) * nb, zh = zb + 3u * unsigned(j) * nb;
            L[j].t = bld(R.tmp, ob, zs);
            L[j].t0 = bld(R.tmp, ob, zs + (e0 ? pb : 0u));
            L[j].t1 = bld(R.tmp, ob1, zs);
            L[j].t2 = bld(R.tmp, ob2, zs);
            L[j].h0 = bld(R.hp, ob, zh);
            L[j].h1 = bld(R.hp, ob, zh + nb);
            L[j].h2 = bld(R.hp, ob, zh + 2u * nb);
        }
        C.q = bld4(R.q, 4u * ob, 4u * zb);
#pragma unroll
        for (int j = 0; j < NS; ++j) L[j].a = bld(R.a, ob, zb + unsigned(j) * nb);
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            L[j].pt = bld(R.pt, ob, zb + unsigned(j) * nb);
            L[j].u = bld(R.u, ob, zb + unsigned(j) * nb);
        }
        C.cs = bld(R.cs, ob, zb);
        C.ct = bld(R.ct, ob, zb);
    }
}

// Thread t -> point (yy, xx) of the widened tile (TY + 3) x (TX + 3).
// MAP 0: row-major. MAP 1: whole rows (a wave's loads stay contiguous), the
// rows of the output region O (yy = 2 .. TY+1) first, then the ring rows that
// only the earlier levels need: yy = 1 (level-B p-step, level-A u-step),
// yy = TY + 2 (level-A u-step), yy = 0 (level-A p-step only). MAP 2: O
// row-major first, then each ring (rows and columns) by the last phase it
// takes part in: E1 \ O, E2 \ E1, E \ E2. Waves made only of ring points
// skip the later phases; MAP 2 skips more but loads the ring columns one
// point per lane.
template <int TY, int TX, int MAP>
__device__ __forceinline__ void sweep_point(int t, int &yy, int &xx)
{
    constexpr int EX = TX + 3;
    if (MAP == 0) {
        yy = t / EX;
        xx = t % EX;
        return;
    }
    if (MAP == 1) {
        const int r = t / EX;
        xx = t % EX;
        yy = r < TY ? r + 2 : (r == TY ? 1 : (r == TY + 1 ? TY + 2 : 0));
        return;
    }
    yy = xx = 0;
    if (t < TY * TX) {
        yy = 2 + t / TX;
        xx = 2 + t % TX;
        return;
    }
    t -= TY * TX;
    if (t < TX + 1 + TY) {  // E1 \ O: row 1, then column 1
        if (t < TX + 1) { yy = 1; xx = 1 + t; } else { yy = 2 + (t - (TX + 1)); xx = 1; }
        return;
    }
    t -= TX + 1 + TY;
    if (t < TX + 2 + TY + 1) {  // E2 \ E1: row TY + 2, then column TX + 2
        if (t < TX + 2) { yy = TY + 2; xx = 1 + t; } else { yy = 1 + (t - (TX + 2)); xx = TX + 2; }
        return;
    }
    t -= TX + 2 + TY + 1;
    if (t < TX + 3 + TY + 2) {  // E \ E2: row 0, then column 0
        if (t < TX + 3) { yy = 0; xx = t; } else { yy = 1 + (t - (TX + 3)); xx = 0; }
    }
}

// NS solves per thread (NS = 2 only in a batch: the pair shares the frames,
// capacities, masks and barriers, and gives each wave two independent chains
// of arithmetic). Solve j of the pair is at element offset j n of every
// per-solve buffer (hp: 3 j n), so the descriptors span NS solves and a
// solve is a uniform byte offset.
template <int TY, int TX, int NT, int MAP, bool BATCH, int NS>
__global__ void __launch_bounds__(NT)
cmf_sweep_kernel(SweepBufs B, const float4 *__restrict__ q4, const float *__restrict__ alpha,
                 const float *__restrict__ cs, const float *__restrict__ ct, const Dims D,
                 const PParams PP, const UParams UP, int zchunk, int levels, int check,
                 int out_set, int *__restrict__ ctrl, float *__restrict__ partials,
                 const SweepBatch SB)
{
    static_assert(NS == 1 || (BATCH && NS == 2), "two solves per thread only in a batch");
    // block -> (first of the NS solves of the batch it serves, tile)
    unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    int ks = 0;
    if (BATCH) {
        const unsigned units = unsigned(SB.ns / NS);
        ks = int(b % units) * NS;
        b /= units;
    }
    // a stopped solve is neither advanced nor reported
    bool act[NS];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        act[j] = !(BATCH ? SB.stop[SB.s0 + ks + j] : ctrl[kStop]);
        any = any || act[j];
    }
    if (!any) return;
    const int64_t nvox = int64_t(D.d0) * D.d1 * D.d2;
    if (BATCH) {
        const int64_t o = nvox * ks;
        B.hp += 3 * o;
        B.tmp += o;
        B.pt += o;
        B.u += o;
        B.hpo += 3 * o;
        B.tmpo += o;
        B.pto += o;
        B.uo += o;
        alpha += o;
    }
    constexpr int EY = TY + 3, EX = TX + 3, NE = EY * EX;
    static_assert(NE <= NT, "one point per thread");
    constexpr bool PST = NT >= 1024;
    // per solve: axis-1 and axis-2 components of the hp planes (the axis-0
    // component is only needed at the thread's own voxel), and level-A tmp
    // of plane z-1
    __shared__ float sA1[NS][NE], sA2[NS][NE];  // level-A hp at plane z
    __shared__ float sT[NS][NE];                // level-A tmp at plane z-1
    __shared__ float sB1[NS][NE], sB2[NS][NE];  // level-B hp at plane z-1

    const int ntx = (D.d2 + TX - 1) / TX, nty = (D.d1 + TY - 1) / TY;
    const int tx = int(b % unsigned(ntx)), ty = int((b / unsigned(ntx)) % unsigned(nty));
    const int tz = int(b / unsigned(ntx * nty));
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * zchunk;
    const int z1 = min(z0 + zchunk, D.d0);
    const int64_t n = nvox;
    const int64_t st0 = int64_t(D.d1) * D.d2;
    const SweepRsrc R{rsrc_of(B.tmp, 4 * n * NS), rsrc_of(B.hp, 12 * n * NS), rsrc_of(q4, 16 * n),
                      rsrc_of(alpha, 4 * n * NS), rsrc_of(B.pt, 4 * n * NS), rsrc_of(B.u, 4 * n * NS),
                      rsrc_of(cs, 4 * n), rsrc_of(ct, 4 * n), rsrc_of(B.hpo, 12 * n * NS),
                      rsrc_of(B.tmpo, 4 * n * NS), rsrc_of(B.pto, 4 * n * NS),
                      rsrc_of(B.uo, 4 * n * NS), B.tmp, B.hp, alpha, B.pt, B.u, cs, ct, q4};
    const unsigned nb = unsigned(4 * n), pb = unsigned(4 * st0), d2b = 4u * unsigned(D.d2);

    // this thread's point of the widened tile; LDS is indexed by the point
    // p = yy EX + xx
    int yy, xx;
    sweep_point<TY, TX, MAP>(int(threadIdx.x), yy, xx);
    const bool live = int(threadIdx.x) < NE;
    const int p = yy * EX + xx;
    const int y = y0 - 2 + yy, x = x0 - 2 + xx;
    const bool inv = live && y >= 0 && y < D.d1 && x >= 0 && x < D.d2;
    const bool inE2 = inv && yy >= 1 && xx >= 1;
    const bool inE1 = inv && yy >= 1 && yy < TY + 2 && xx >= 1 && xx < TX + 2;
    const bool inO = inv && yy >= 2 && yy < TY + 2 && xx >= 2 && xx < TX + 2;
    const unsigned f = (y < D.d1 - 1 ? 16u : 0u) | (x < D.d2 - 1 ? 32u : 0u) |
                       (y > 0 ? 64u : 0u) | (x > 0 ? 128u : 0u);
    const unsigned ob = inv ? 4u * unsigned(y * D.d2 + x) : 0u;  // in-plane byte offset

    const bool two = levels == 2;
    const int zs = max(z0 - (two ? 2 : 1), 0);
    const int zend = two ? z1 : z1 - 1;
    auto p_on = [&](int z) { return inv && z < D.d0; };
    auto u_on = [&](int z) { return (two ? inE2 : inO) && z < D.d0 && z >= (two ? z0 - 1 : z0); };

    // values carried from step z-1 to step z
    float hA1_0[NS], hA1_1[NS], hA1_2[NS];  // level-A hp at z-1
    float hB2_0[NS];                        // level-B hp (axis 0) at z-2
    float ptAp[NS], uAp[NS], tAp[NS];       // level-A pt, u, tmp at z-1
    float aP[NS];
    float err_acc[NS];
    float csP = 0.f, ctP = 0.f;
    float vP0 = 1.f, vP1 = 0.f, vP2 = 0.f;  // frame axis at z-1
    SweepIn L[NS];
    SweepInShared LC;
    LC.q = make_float4(1.f, 0.f, 0.f, 0.f);
    LC.cs = LC.ct = 0.f;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        hA1_0[j] = hA1_1[j] = hA1_2[j] = hB2_0[j] = 0.f;
        ptAp[j] = uAp[j] = tAp[j] = aP[j] = err_acc[j] = 0.f;
        L[j].t = L[j].t0 = L[j].t1 = L[j].t2 = L[j].h0 = L[j].h1 = L[j].h2 = 0.f;
        L[j].a = L[j].pt = L[j].u = 0.f;
    }
    // the forward neighbours' in-plane offsets, the point itself on the far
    // faces and outside the volume (where f describes no point of it)
    const unsigned ob1 = (inv && (f & 16u)) ? ob + d2b : ob;
    const unsigned ob2 = (inv && (f & 32u)) ? ob + 4u : ob;
    auto load = [&](const int z) {
        if constexpr (NS == 1) {
            const int zc = min(z, D.d0 - 1);  // a plane past the volume is not used
            sweep_load_all<NS, PST>(L, LC, R, nb, pb, unsigned(zc) * pb, ob, ob1, ob2,
                                    zc < D.d0 - 1);
        } else {
            sweep_load<NS>(L, LC, R, nb, pb, unsigned(z) * pb, ob, ob1, ob2, p_on(z), z < D.d0 - 1);
        }
    };
    if (zs <= zend) load(zs);
    // one plane step; the loop below takes two per turn, so that the values
    // carried between steps alternate between two register sets instead of
    // being copied back at every step
    auto step = [&](const int z) {
        const bool ph1 = p_on(z), ph2 = u_on(z);
        const bool ph3 = two && inE1 && z >= max(z0, 1);
        const bool ph4 = two && inO && z - 1 >= z0 && z - 1 < z1;
        const bool e0 = z < D.d0 - 1, f0 = z > 0;            // plane z
        const bool e0b = z - 1 < D.d0 - 1, f0b = z - 1 > 0;  // plane z-1
        const unsigned zb = unsigned(z) * pb;  // plane z, bytes
        // 1. level-A p-step at plane z on the widened tile
        // (lanes outside a phase's region keep values no valid lane reads)
        float hA0_0[NS], hA0_1[NS], hA0_2[NS];
        // the frame axis of plane z, formed once for both levels and every solve
        float vN0, vN1, vN2;
        mh::frame_axis(LC.q, vN0, vN1, vN2);
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            hA0_0[j] = L[j].h0;
            hA0_1[j] = L[j].h1;
            hA0_2[j] = L[j].h2;
        }
        if (ph1) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const float g0 = e0 ? (L[j].t0 - L[j].t) * PP.ss0 : 0.f;
                const float g1 = (f & 16u) ? (L[j].t1 - L[j].t) * PP.ss1 : 0.f;
                const float g2 = (f & 32u) ? (L[j].t2 - L[j].t) * PP.ss2 : 0.f;
                mh::cmf_p_voxel_axi(hA0_0[j], hA0_1[j], hA0_2[j], g0, g1, g2, vN0, vN1, vN2, PP.A,
                                    L[j].a);
            }
            if (!two && inO && z >= z0) {
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    if (NS > 1 && !act[j]) continue;
                    const unsigned zh = zb + 3u * unsigned(j) * nb;
                    sweep_store<PST>(R.hpo, B.hpo, ob, zh, hA0_0[j]);
                    sweep_store<PST>(R.hpo, B.hpo, ob, zh + nb, hA0_1[j]);
                    sweep_store<PST>(R.hpo, B.hpo, ob, zh + 2u * nb, hA0_2[j]);
                }
            }
        }
        if (live) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                sA1[j][p] = hA0_1[j];
                sA2[j][p] = hA0_2[j];
            }
        }
        __syncthreads();
        // 2. level-A u-step at plane z (stores when it is the only level)
        float tA0[NS], ptA0[NS], uA0[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            tA0[j] = L[j].t;
            ptA0[j] = L[j].pt;
            uA0[j] = L[j].u;
        }
        if (ph2) {
            float err[NS];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                // E2 holds p - EX and p - 1; the faces mask what they read
                const float n1 = sA1[j][p - EX], n2 = sA2[j][p - 1];
                const float x0v = (e0 ? hA0_0[j] : 0.f) - (f0 ? hA1_0[j] : 0.f);
                const float x1v = ((f & 16u) ? hA0_1[j] : 0.f) - ((f & 64u) ? n1 : 0.f);
                const float x2v = ((f & 32u) ? hA0_2[j] : 0.f) - ((f & 128u) ? n2 : 0.f);
                const float div = mh::cmf_div(x0v, x1v, x2v, UP.rs0, UP.rs1, UP.rs2);
                err[j] = mh::cmf_u_voxel(div, ptA0[j], uA0[j], tA0[j], LC.cs, LC.ct, UP.rcc,
                                         UP.inv_rcc, 1.f, 0.f);
            }
            if (!two) {
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    if (check) err_acc[j] += fabsf(err[j]);
                    if (NS > 1 && !act[j]) continue;
                    const unsigned zj = zb + unsigned(j) * nb;
                    sweep_store<PST>(R.pto, B.pto, ob, zj, ptA0[j]);
                    sweep_store<PST>(R.uo, B.uo, ob, zj, uA0[j]);
                    sweep_store<PST>(R.tmpo, B.tmpo, ob, zj, tA0[j]);
                }
            }
        }
        // constants of plane z, needed at plane z-1's turn in the next step
        float aN[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) aN[j] = L[j].a;
        const float csN = LC.cs, ctN = LC.ct;
        // inputs of the next step, in flight while the level-B work runs
        if (z + 1 <= zend) load(z + 1);
        // 3. level-B p-step at plane z-1 on E1
        float hB1_0[NS], hB1_1[NS], hB1_2[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            hB1_0[j] = hA1_0[j];
            hB1_1[j] = hA1_1[j];
            hB1_2[j] = hA1_2[j];
        }
        if (ph3) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                // E holds p + EX and p + 1 of every point of E1
                const float t1 = sT[j][p + EX], t2 = sT[j][p + 1];
                const float g0 = e0b ? (tA0[j] - tAp[j]) * PP.ss0 : 0.f;
                const float g1 = (f & 16u) ? (t1 - tAp[j]) * PP.ss1 : 0.f;
                const float g2 = (f & 32u) ? (t2 - tAp[j]) * PP.ss2 : 0.f;
                mh::cmf_p_voxel_axi(hB1_0[j], hB1_1[j], hB1_2[j], g0, g1, g2, vP0, vP1, vP2, PP.A,
                                    aP[j]);
            }
        }
        if (live) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                sB1[j][p] = hB1_1[j];
                sB2[j][p] = hB1_2[j];
            }
        }
        __syncthreads();
        // 4. level-B u-step at plane z-1 on O; store iteration t + 2
        if (ph4) {
            const unsigned zp = zb - pb;  // plane z-1
            float pti[NS], ui[NS], ti[NS], err[NS];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const float n1 = sB1[j][p - EX], n2 = sB2[j][p - 1];  // O: inside E1
                const float x0v = (e0b ? hB1_0[j] : 0.f) - (f0b ? hB2_0[j] : 0.f);
                const float x1v = ((f & 16u) ? hB1_1[j] : 0.f) - ((f & 64u) ? n1 : 0.f);
                const float x2v = ((f & 32u) ? hB1_2[j] : 0.f) - ((f & 128u) ? n2 : 0.f);
                const float div = mh::cmf_div(x0v, x1v, x2v, UP.rs0, UP.rs1, UP.rs2);
                pti[j] = ptAp[j];
                ui[j] = uAp[j];
                err[j] = mh::cmf_u_voxel(div, pti[j], ui[j], ti[j], csP, ctP, UP.rcc, UP.inv_rcc,
                                         1.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (check) err_acc[j] += fabsf(err[j]);
                if (NS > 1 && !act[j]) continue;
                const unsigned zj = zp + unsigned(j) * nb, zh = zp + 3u * unsigned(j) * nb;
                sweep_store<PST>(R.hpo, B.hpo, ob, zh, hB1_0[j]);
                sweep_store<PST>(R.hpo, B.hpo, ob, zh + nb, hB1_1[j]);
                sweep_store<PST>(R.hpo, B.hpo, ob, zh + 2u * nb, hB1_2[j]);
                sweep_store<PST>(R.pto, B.pto, ob, zj, pti[j]);
                sweep_store<PST>(R.uo, B.uo, ob, zj, ui[j]);
                sweep_store<PST>(R.tmpo, B.tmpo, ob, zj, ti[j]);
            }
        }
        // level-A tmp of plane z for the next step's level-B p-step (every read
        // of plane z-1's values finished before the barrier above)
        if (live) {
#pragma unroll
            for (int j = 0; j < NS; ++j) sT[j][p] = tA0[j];
        }
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            hB2_0[j] = hB1_0[j];
            hA1_0[j] = hA0_0[j];
            hA1_1[j] = hA0_1[j];
            hA1_2[j] = hA0_2[j];
            ptAp[j] = ptA0[j];
            uAp[j] = uA0[j];
            tAp[j] = tA0[j];
            aP[j] = aN[j];
        }
        vP0 = vN0;
        vP1 = vN1;
        vP2 = vN2;
        csP = csN;
        ctP = ctN;
    };
    for (int z = zs; z <= zend; z += 2) {
        step(z);
        if (z + 1 > zend) break;  // uniform: zs and zend are the work-group's
        step(z + 1);
    }
    if (check) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const float v = mh::block_reduce<SumOp>(err_acc[j]);
            if (threadIdx.x == 0 && (NS == 1 || act[j])) {
                if (BATCH) partials[int64_t(SB.s0 + ks + j) * SB.nper + b] = v;
                else partials[blockIdx.x] = v;
            }
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (NS > 1 && !act[j]) continue;
            if (BATCH && b == 0) SB.parity[SB.s0 + ks + j] = out_set;
            else if (!BATCH && blockIdx.x == 0) ctrl[kParity] = out_set;
        }
    }
}

// ------------------------------------------------------ L-level sweep ---
// L iterations (levels 0 .. L-1) in one pass over the state, 1024 threads,
// NS solves per thread (NS = 2 only in a batch: the pair shares the frame,
// cs, ct, their loads and the barriers; alpha is per solve): the two-level scheme with more
// levels. The widened tile is (TY + 2L - 1) x (TX + 2L - 1) points and the
// output tile starts at (L, L). At step z level l works at plane z - l:
//   its p-step on rows and columns l .. T + 2L - 2 - l, from level l-1's hp
//     at that plane (the loaded state for level 0) and level l-1's tmp there
//     and at the plane above (+1 in-plane neighbours from LDS),
//   a barrier (its hp along axes 1 and 2 is in LDS),
//   its u-step on rows and columns l + 1 .. T + 2L - 2 - l, from its hp and
//     the -1 in-plane neighbours' (LDS), its hp along axis 0 at the plane
//     below (carried), and level l-1's pt and u there;
// the last level's u-step covers the output tile and stores iteration t + L.
// Each level's tmp goes to LDS at the end of the step, for the next level's
// p-step at the next step. Per plane the state is read and written once per
// L iterations; each level's carried values and plane constants (alpha,
// frame axis, cs, ct) shift by one level per step. The per-voxel arithmetic
// is the shared helpers', so a pass equals L launches of the plain kernels
// bit for bit.
template <int TY, int TX, int L, bool BATCH, int NS>
__global__ void __launch_bounds__(1024)
cmf_sweepL_kernel(SweepBufs B, const float4 *__restrict__ q4, const float *__restrict__ alpha,
                  const float *__restrict__ cs, const float *__restrict__ ct, const Dims D,
                  const PParams PP, const UParams UP, int zchunk, int check, int out_set,
                  int *__restrict__ ctrl, float *__restrict__ partials, const SweepBatch SB)
{
    static_assert(L >= 2 && L <= 6, "2 to 6 levels");
    static_assert(NS == 1 || (BATCH && NS == 2), "two solves per thread only in a batch");
    unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    int ks = 0;
    if (BATCH) {
        const unsigned units = unsigned(SB.ns / NS);
        ks = int(b % units) * NS;
        b /= units;
    }
    // a stopped solve is neither advanced nor reported
    bool act[NS];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        act[j] = !(BATCH ? SB.stop[SB.s0 + ks + j] : ctrl[kStop]);
        any = any || act[j];
    }
    if (!any) return;
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    if (BATCH) {
        const int64_t o = n * ks;
        B.hp += 3 * o;
        B.tmp += o;
        B.pt += o;
        B.u += o;
        B.hpo += 3 * o;
        B.tmpo += o;
        B.pto += o;
        B.uo += o;
        alpha += o;
    }
    constexpr int W = 2 * L - 1;  // halo width of the widened tile
    constexpr int EY = TY + W, EX = TX + W, NE = EY * EX;
    static_assert(NE <= 1024, "one point per thread");
    __shared__ float sH1[L][NS][NE], sH2[L][NS][NE];  // hp of each level along axes 1, 2
    __shared__ float sT[L - 1][NS][NE];                // tmp of levels 0 .. L-2

    const int ntx = (D.d2 + TX - 1) / TX, nty = (D.d1 + TY - 1) / TY;
    const int tx = int(b % unsigned(ntx)), ty = int((b / unsigned(ntx)) % unsigned(nty));
    const int tz = int(b / unsigned(ntx * nty));
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * zchunk;
    const int z1 = min(z0 + zchunk, D.d0);
    const int64_t st0 = int64_t(D.d1) * D.d2;
    const SweepRsrc R{rsrc_of(B.tmp, 4 * n * NS), rsrc_of(B.hp, 12 * n * NS), rsrc_of(q4, 16 * n),
                      rsrc_of(alpha, 4 * n * NS), rsrc_of(B.pt, 4 * n * NS), rsrc_of(B.u, 4 * n * NS),
                      rsrc_of(cs, 4 * n), rsrc_of(ct, 4 * n), rsrc_of(B.hpo, 12 * n * NS),
                      rsrc_of(B.tmpo, 4 * n * NS), rsrc_of(B.pto, 4 * n * NS),
                      rsrc_of(B.uo, 4 * n * NS), B.tmp, B.hp, alpha, B.pt, B.u, cs, ct, q4};
    const unsigned nb = unsigned(4 * n), pb = unsigned(4 * st0), d2b = 4u * unsigned(D.d2);

    const int t = int(threadIdx.x);
    const int yy = t / EX, xx = t % EX;
    const bool live = t < NE;
    const int p = yy * EX + xx;
    const int y = y0 - L + yy, x = x0 - L + xx;
    const bool inv = live && y >= 0 && y < D.d1 && x >= 0 && x < D.d2;
    // the in-plane regions of level l: p-step rows/columns l .. hi_l - 1,
    // u-step l + 1 .. hi_l - 1, with hi_l = T + 2L - 1 - l
    bool inP[L], inU[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        inP[l] = inv && yy >= l && yy < TY + W - l && xx >= l && xx < TX + W - l;
        inU[l] = inv && yy >= l + 1 && yy < TY + W - l && xx >= l + 1 && xx < TX + W - l;
    }
    const unsigned f = (y < D.d1 - 1 ? 16u : 0u) | (x < D.d2 - 1 ? 32u : 0u) |
                       (y > 0 ? 64u : 0u) | (x > 0 ? 128u : 0u);
    const unsigned ob = inv ? 4u * unsigned(y * D.d2 + x) : 0u;
    const unsigned ob1 = (inv && (f & 16u)) ? ob + d2b : ob;
    const unsigned ob2 = (inv && (f & 32u)) ? ob + 4u : ob;

    const int zs = max(z0 - L, 0), zend = z1 + L - 2;
    // The loop over steps is unrolled U = lcm(2, L) times and every carried
    // value sits in a ring indexed by the step's phase PH (compile-time), so
    // nothing is copied from step to step:
    //   per level l and solve j, hp, tmp, pt, u at plane z - l in slot PH & 1
    //   (written this step) and at plane z - 1 - l in slot (PH + 1) & 1 (the
    //   previous step's);
    //   the plane constants (alpha per solve, cs, ct, frame axis) of plane p
    //   in slot p mod L: plane z - l in slot (PH - l) mod L.
    constexpr int U = L % 2 == 0 ? L : 2 * L;
    float h0[2][L][NS], h1[2][L][NS], h2[2][L][NS], tp[2][L][NS], ptp[2][L][NS], up[2][L][NS];
    float ac[L][NS], csc[L], ctc[L], v0[L], v1[L], v2[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
#pragma unroll
            for (int k = 0; k < 2; ++k)
                h0[k][l][j] = h1[k][l][j] = h2[k][l][j] = tp[k][l][j] = ptp[k][l][j] = up[k][l][j] = 0.f;
            ac[l][j] = 0.f;
        }
        csc[l] = ctc[l] = 0.f;
        v0[l] = 1.f;
        v1[l] = v2[l] = 0.f;
    }
    float err_acc[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) err_acc[j] = 0.f;
    SweepIn Li[NS];
    SweepInShared LC;
    LC.q = make_float4(1.f, 0.f, 0.f, 0.f);
    LC.cs = LC.ct = 0.f;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        Li[j].t = Li[j].t0 = Li[j].t1 = Li[j].t2 = Li[j].h0 = Li[j].h1 = Li[j].h2 = 0.f;
        Li[j].a = Li[j].pt = Li[j].u = 0.f;
    }
    auto load = [&](const int z) {
        const int zc = min(z, D.d0 - 1);
        sweep_load_all<NS, true>(Li, LC, R, nb, pb, unsigned(zc) * pb, ob, ob1, ob2, zc < D.d0 - 1);
    };
    if (zs < D.d0) load(zs);
    auto step = [&](auto phase, const int z) {
        constexpr int PH = decltype(phase)::value;
        constexpr int C = PH & 1, P = C ^ 1;   // this step's and the previous step's slots
        // plane constants of plane z (level 0), from the loads, into slot PH mod L
        // (its last reader, level L-1 at plane z - L, ran the step before)
        mh::frame_axis(LC.q, v0[PH % L], v1[PH % L], v2[PH % L]);
#pragma unroll
        for (int j = 0; j < NS; ++j) ac[PH % L][j] = Li[j].a;
        csc[PH % L] = LC.cs;
        ctc[PH % L] = LC.ct;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int w = z - l;
            const int cl = (PH - l + L * U) % L;   // plane w's constants
            const bool pon = inP[l] && w >= max(z0 - (L - l), 0) && w < min(z1 + (L - 1 - l), D.d0);
            const bool uon = inU[l] && w >= max(z0 - (L - 1 - l), 0) && w < min(z1 + (L - 1 - l), D.d0);
            const bool e0 = w < D.d0 - 1, f0 = w > 0;
            const float c_s = csc[cl], c_t = ctc[cl];
            const float w0 = v0[cl], w1 = v1[cl], w2 = v2[cl];
            // p-step: from level l-1's hp at plane w (previous step) or the loads
            float g0h[NS], g1h[NS], g2h[NS];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                g0h[j] = l == 0 ? Li[j].h0 : h0[P][l - 1][j];
                g1h[j] = l == 0 ? Li[j].h1 : h1[P][l - 1][j];
                g2h[j] = l == 0 ? Li[j].h2 : h2[P][l - 1][j];
            }
            if (pon) {
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    float g0, g1, g2;
                    if (l == 0) {
                        g0 = e0 ? (Li[j].t0 - Li[j].t) * PP.ss0 : 0.f;
                        g1 = (f & 16u) ? (Li[j].t1 - Li[j].t) * PP.ss1 : 0.f;
                        g2 = (f & 32u) ? (Li[j].t2 - Li[j].t) * PP.ss2 : 0.f;
                    } else {
                        // level l-1's tmp: at plane w (previous step, and its +1
                        // in-plane neighbours in LDS) and at w+1 (this step)
                        const float tc = tp[P][l - 1][j];
                        const float t1 = sT[l - 1][j][p + EX], t2 = sT[l - 1][j][p + 1];
                        g0 = e0 ? (tp[C][l - 1][j] - tc) * PP.ss0 : 0.f;
                        g1 = (f & 16u) ? (t1 - tc) * PP.ss1 : 0.f;
                        g2 = (f & 32u) ? (t2 - tc) * PP.ss2 : 0.f;
                    }
                    mh::cmf_p_voxel_axi(g0h[j], g1h[j], g2h[j], g0, g1, g2, w0, w1, w2, PP.A, ac[cl][j]);
                }
            }
            if (live) {
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    sH1[l][j][p] = g1h[j];
                    sH2[l][j][p] = g2h[j];
                }
            }
            __syncthreads();
            // u-step: from level l-1's pt, u at plane w (previous step) or the loads
            float tt[NS], ptt[NS], uu[NS];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                tt[j] = l == 0 ? Li[j].t : 0.f;
                ptt[j] = l == 0 ? Li[j].pt : ptp[P][l - 1][j];
                uu[j] = l == 0 ? Li[j].u : up[P][l - 1][j];
            }
            if (uon) {
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    const float n1 = sH1[l][j][p - EX], n2 = sH2[l][j][p - 1];
                    const float x0v = (e0 ? g0h[j] : 0.f) - (f0 ? h0[P][l][j] : 0.f);
                    const float x1v = ((f & 16u) ? g1h[j] : 0.f) - ((f & 64u) ? n1 : 0.f);
                    const float x2v = ((f & 32u) ? g2h[j] : 0.f) - ((f & 128u) ? n2 : 0.f);
                    const float div = mh::cmf_div(x0v, x1v, x2v, UP.rs0, UP.rs1, UP.rs2);
                    const float err = mh::cmf_u_voxel(div, ptt[j], uu[j], tt[j], c_s, c_t, UP.rcc,
                                                      UP.inv_rcc, 1.f, 0.f);
                    if (l == L - 1) {
                        if (check) err_acc[j] += fabsf(err);
                        if (NS > 1 && !act[j]) continue;
                        const unsigned zp = unsigned(w) * pb, zj = zp + unsigned(j) * nb,
                                       zh = zp + 3u * unsigned(j) * nb;
                        pst(B.hpo, ob, zh, g0h[j]);
                        pst(B.hpo, ob, zh + nb, g1h[j]);
                        pst(B.hpo, ob, zh + 2u * nb, g2h[j]);
                        pst(B.pto, ob, zj, ptt[j]);
                        pst(B.uo, ob, zj, uu[j]);
                        pst(B.tmpo, ob, zj, tt[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                h0[C][l][j] = g0h[j];
                h1[C][l][j] = g1h[j];
                h2[C][l][j] = g2h[j];
                tp[C][l][j] = tt[j];
                ptp[C][l][j] = ptt[j];
                up[C][l][j] = uu[j];
            }
            // the next plane's inputs, in flight during the later levels
            if (l == 0 && z + 1 <= zend && z + 1 < D.d0) load(z + 1);
        }
        // tmp of levels 0 .. L-2 for the next step's p-steps (every read of
        // the previous ones finished before the barriers above)
        if (live) {
#pragma unroll
            for (int l = 0; l < L - 1; ++l)
#pragma unroll
                for (int j = 0; j < NS; ++j) sT[l][j][p] = tp[C][l][j];
        }
    };
    for (int z = zs; z <= zend; z += U) {
        // U steps with compile-time phases; z - zs is a multiple of U here
        const bool go = mh::unrolled_while<U>([&](auto ph) {
            constexpr int K = decltype(ph)::value;
            if (z + K > zend) return false;   // uniform: zs and zend are the work-group's
            step(ph, z + K);
            return true;
        });
        if (!go) break;
    }
    if (check) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const float v = mh::block_reduce<SumOp>(err_acc[j]);
            if (threadIdx.x == 0 && (NS == 1 || act[j])) {
                if (BATCH) partials[int64_t(SB.s0 + ks + j) * SB.nper + b] = v;
                else partials[blockIdx.x] = v;
            }
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (NS > 1 && !act[j]) continue;
            if (BATCH && b == 0) SB.parity[SB.s0 + ks + j] = out_set;
            else if (!BATCH && blockIdx.x == 0) ctrl[kParity] = out_set;
        }
    }
}

// batch: NS solves per thread, grid = work-groups per solve x batch->ns / NS
template <int TY, int TX, int NT, int MAP>
void launch_sweep_t(unsigned grid, hipStream_t s, const SweepBufs &SBf, const float4 *q4,
                    const float *alpha, const float *cs, const float *ct, const Dims &D,
                    const PParams &PP, const UParams &UP, int zchunk, int lv, int chk, int out_set,
                    int *ctrl, float *partials, const SweepBatch *batch, int ns_thread)
{
    if (batch && ns_thread == 2)
        hipLaunchKernelGGL((cmf_sweep_kernel<TY, TX, NT, MAP, true, 2>), dim3(grid), dim3(NT), 0, s,
                           SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl,
                           partials, *batch);
    else if (batch)
        hipLaunchKernelGGL((cmf_sweep_kernel<TY, TX, NT, MAP, true, 1>), dim3(grid), dim3(NT), 0, s,
                           SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl,
                           partials, *batch);
    else
        hipLaunchKernelGGL((cmf_sweep_kernel<TY, TX, NT, MAP, false, 1>), dim3(grid), dim3(NT), 0, s,
                           SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl,
                           partials, SweepBatch{});
}

// L levels: lv must be L (shorter passes take a shape of fewer levels)
template <int TY, int TX, int L>
void launch_sweepL_t(unsigned grid, hipStream_t s, const SweepBufs &SBf, const float4 *q4,
                     const float *alpha, const float *cs, const float *ct, const Dims &D,
                     const PParams &PP, const UParams &UP, int zchunk, int chk, int out_set, int *ctrl,
                     float *partials, const SweepBatch *batch, int ns_thread)
{
    if constexpr (L <= 3) {  // pairs only up to three levels (sweep_pairs)
        if (batch && ns_thread == 2) {
            hipLaunchKernelGGL((cmf_sweepL_kernel<TY, TX, L, true, 2>), dim3(grid), dim3(1024), 0, s, SBf,
                               q4, alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, *batch);
            return;
        }
    }
    if (batch)
        hipLaunchKernelGGL((cmf_sweepL_kernel<TY, TX, L, true, 1>), dim3(grid), dim3(1024), 0, s, SBf, q4,
                           alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, *batch);
    else
        hipLaunchKernelGGL((cmf_sweepL_kernel<TY, TX, L, false, 1>), dim3(grid), dim3(1024), 0, s, SBf, q4,
                           alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, SweepBatch{});
}

// thread-to-point mapping of each tile shape (sweep_point)
constexpr int kMap0 = 2, kMap1 = 1, kMap2 = 2, kMap3 = 2, kMap4 = 0;

template <int TY, int TX, int NT, int MAP>
int sweep_occupancy_t(int &blocks)
{
    return int(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, reinterpret_cast<const void *>(&cmf_sweep_kernel<TY, TX, NT, MAP, false, 1>), NT, 0));
}
template <int TY, int TX, int L>
int sweepL_occupancy_t(int &blocks)
{
    return int(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, reinterpret_cast<const void *>(&cmf_sweepL_kernel<TY, TX, L, false, 1>), 1024, 0));
}

// Work-groups of one sweep launch with this tile shape that the device holds
// at once (CU count x resident work-groups per CU), cached in the context.
int sweep_slots(mh_ctx *ctx, int shape_ix);

// The plane chunk of a tile shape that a launch's cost model favours:
// work-groups run in ceil(groups / slots) waves, each group sweeping its chunk
// plus the planes below and above it that the levels recompute (two for two
// levels, four for three), so the launch takes about waves x (chunk + 2 L - 2)
// plane steps. Chunks giving more than max_groups work-groups are skipped.
int sweep_model_zchunk(mh_ctx *ctx, int shape_ix, const Dims &D, int64_t max_groups)
{
    const SweepShape &S = kSweepShapes[shape_ix];
    const int64_t tiles = int64_t((D.d2 + S.tx - 1) / S.tx) * ((D.d1 + S.ty - 1) / S.ty);
    const int64_t slots = sweep_slots(ctx, shape_ix);
    int best = std::min(16, D.d0);
    int64_t best_cost = -1;
    for (int z = std::min(4, D.d0); z <= std::min(64, D.d0); ++z) {
        const int64_t groups = tiles * ((D.d0 + z - 1) / z);
        if (groups > max_groups) continue;
        const int64_t cost = ((groups + slots - 1) / slots) * (z + 2 * S.levels - 2);
        if (best_cost < 0 || cost < best_cost || (cost == best_cost && z > best)) {
            best_cost = cost;
            best = z;
        }
    }
    return best;
}

int sweep_slots(mh_ctx *ctx, int shape_ix)
{
    int &slots = ctx->sweep_slots[size_t(shape_ix)];
    if (slots > 0) return slots;
    int cus = 0, per = 0, rc = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
        cus = 0;
    switch (shape_ix) {
    case 1: rc = sweep_occupancy_t<11, 32, 512, kMap1>(per); break;
    case 2: rc = sweep_occupancy_t<4, 64, 512, kMap2>(per); break;
    case 3: rc = sweep_occupancy_t<26, 32, 1024, kMap3>(per); break;
    case 4: rc = sweep_occupancy_t<13, 61, 1024, kMap4>(per); break;
    case 5: rc = sweepL_occupancy_t<11, 59, 3>(per); break;
    case 6: rc = sweepL_occupancy_t<27, 27, 3>(per); break;
    case 7: rc = sweepL_occupancy_t<9, 57, 4>(per); break;
    case 8: rc = sweepL_occupancy_t<25, 25, 4>(per); break;
    default: rc = sweep_occupancy_t<12, 64, 1024, kMap0>(per); break;
    }
    (void)hipGetLastError();
    slots = (rc == 0 && cus > 0 && per > 0) ? cus * per : 256;
    return slots;
}

// grid: work-groups per solve times the number of solves (batch->ns)
void launch_sweep(int shape_ix, unsigned grid, hipStream_t s, const SweepBufs &SBf,
                  const float4 *q4, const float *alpha, const float *cs, const float *ct,
                  const Dims &D, const PParams &PP, const UParams &UP, int zchunk, int lv, int chk,
                  int out_set, int *ctrl, float *partials, const SweepBatch *batch = nullptr,
                  int ns_thread = 1)
{
    switch (shape_ix) {
    case 1: launch_sweep_t<11, 32, 512, kMap1>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 2: launch_sweep_t<4, 64, 512, kMap2>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 3: launch_sweep_t<26, 32, 1024, kMap3>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 4: launch_sweep_t<13, 61, 1024, kMap4>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 5: launch_sweepL_t<11, 59, 3>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 6: launch_sweepL_t<27, 27, 3>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 7: launch_sweepL_t<9, 57, 4>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, batch, ns_thread); break;
    case 8: launch_sweepL_t<25, 25, 4>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, chk, out_set, ctrl, partials, batch, ns_thread); break;
    default: launch_sweep_t<12, 64, 1024, kMap0>(grid, s, SBf, q4, alpha, cs, ct, D, PP, UP, zchunk, lv, chk, out_set, ctrl, partials, batch, ns_thread); break;
    }
}

// alpha of the solves of a group: w exp(-|grad I|^2 k), the expression the
// batched p-step forms in registers
__global__ void __launch_bounds__(mh::kBlock)
cmf_group_alpha_kernel(const float *__restrict__ gms, int64_t n, const SolveParams SP,
                       float *__restrict__ alpha)
{
    const int k = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) alpha[n * k + i] = SP.w[k] * expf(-gms[i] * SP.k[k]);
}

// u of every solve of a batched group whose last pass wrote the group buffers
__global__ void __launch_bounds__(mh::kBlock)
cmf_batch_sweep_final_kernel(float *__restrict__ u, const float *__restrict__ u2, int64_t n, int s0,
                             const int *__restrict__ parity)
{
    const int k = blockIdx.y;
    if (parity[s0 + k] != 1) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) u[n * (s0 + k) + i] = u2[n * k + i];
}

// u of the buffer set that holds the final state, copied to the caller's u
// when the last pass wrote the second set.
__global__ void __launch_bounds__(mh::kBlock)
cmf_sweep_final_kernel(float *__restrict__ u, const float *__restrict__ u2, int64_t n,
                       const int *__restrict__ ctrl)
{
    if (ctrl[kParity] != 1) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) u[i] = u2[i];
}

struct Workspace {
    int *ctrl;
    float *crit_log;
    float *partials;
    float *hp;
    float *pt;
    float *tmp;
    float4 *q4;
    // second state set of the multi-level sweep (ping-pong)
    float *hp2, *pt2, *tmp2, *u2;
};

int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

int64_t layout(int64_t n, char *base, Workspace *w)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    Workspace ws;
    ws.ctrl = reinterpret_cast<int *>(take(sizeof(int) * kCtrlWords));
    ws.crit_log = reinterpret_cast<float *>(take(sizeof(float) * kCritLogMax));
    ws.partials = reinterpret_cast<float *>(take(sizeof(float) * (mh::blocks_for(n) + 64)));
    ws.hp = reinterpret_cast<float *>(take(sizeof(float) * 3 * n));
    ws.pt = reinterpret_cast<float *>(take(sizeof(float) * n));
    ws.tmp = reinterpret_cast<float *>(take(sizeof(float) * n));
    ws.q4 = reinterpret_cast<float4 *>(take(sizeof(float4) * n));
    ws.hp2 = reinterpret_cast<float *>(take(sizeof(float) * 3 * n));
    ws.pt2 = reinterpret_cast<float *>(take(sizeof(float) * n));
    ws.tmp2 = reinterpret_cast<float *>(take(sizeof(float) * n));
    ws.u2 = reinterpret_cast<float *>(take(sizeof(float) * n));
    if (w) *w = ws;
    return off;
}

__global__ void zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

namespace {

// One solve on workspace W, advanced in steps. An iteration runs either as
// the plain kernels (in place on the current buffer set) or inside a
// multi-level sweep (current set -> the other set); both compute the same
// values, so a solve may mix them, and the final u is in the set the last
// sweep wrote (ctrl[kParity]).
struct Runner {
    mh_ctx *ctx;
    float *u;
    const float *eigs, *alpha, *cs, *ct;
    const uint8_t *labels;
    Dims D;
    int64_t n;
    unsigned nb;
    const mh_cmf_params *prm;
    Workspace W;
    hipStream_t s;
    PParams PP;
    UParams UP;
    bool sup, sweep_elig;
    int cur = 0, it = 0, limit = 0;
    bool swept = false;
    bool graphs_allowed = true;  // false: never capture (no stream or graph is created)

    Runner(mh_ctx *c, float *u_, const float *e, const float *a, const float *cs_, const float *ct_,
           const uint8_t *l, Dims d, int64_t n_, const mh_cmf_params *p, const Workspace &w,
           hipStream_t st)
        : ctx(c), u(u_), eigs(e), alpha(a), cs(cs_), ct(ct_), labels(l), D(d), n(n_),
          nb(mh::blocks_for(n_)), prm(p), W(w), s(st), PP(p_params(p)),
          UP{p->rcc, float(1.0 / double(p->rcc)), p->rs[0], p->rs[1], p->rs[2], p->error_bound,
             float(1.0 / (double(n_) * double(p->rcc)))},
          sup(l != nullptr),
          sweep_elig(p->const_evals && l == nullptr && n_ < kSweepMaxVoxels &&
                     mh::axi_params(p).on),
          limit(p->num_iter)
    {
    }

    int init()
    {
        hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, W.ctrl, int(kCtrlWords));
        const int es = prm->const_evals ? 3 : 6, eo = prm->const_evals ? 0 : 3;
        hipLaunchKernelGGL(cmf_init_kernel, dim3(nb), dim3(mh::kBlock), 0, s, W.q4, eigs, es, eo,
                           W.hp, W.pt, W.tmp, u, cs, ct, n, prm->rcc);
        MH_CHECK_LAUNCH(ctx, "cmf init");
        return MH_OK;
    }

    int64_t sweep_grid_of(const SweepMethod &m) const
    {
        const SweepShape &S = kSweepShapes[m.shape];
        return int64_t((D.d2 + S.tx - 1) / S.tx) * ((D.d1 + S.ty - 1) / S.ty) *
               ((D.d0 + m.zchunk - 1) / m.zchunk);
    }
    bool usable(const SweepMethod &m) const
    {
        return m.shape < 0 || (sweep_elig && sweep_grid_of(m) <= int64_t(nb) + 64);
    }

    int model_zchunk(int shape_ix) const { return sweep_model_zchunk(ctx, shape_ix, D, int64_t(nb) + 64); }
    // the method of a pass of lv levels under method m
    SweepMethod pass_method(const SweepMethod &m, int lv) const
    {
        const int levels = kSweepShapes[m.shape].levels;
        if (lv == levels || (levels == 2 && lv < 2)) return m;
        const int k = lv <= 2 ? 0 : kShape3;
        return SweepMethod{k, model_zchunk(k)};
    }

    // one iteration (plain kernels) or one pass of up to L (sweep) that
    // does not go past `limit`; returns the iterations advanced
    int step(const SweepMethod &m)
    {
        if (m.shape < 0) {
            float *hp = cur ? W.hp2 : W.hp, *tmp = cur ? W.tmp2 : W.tmp;
            float *pt = cur ? W.pt2 : W.pt, *uu = cur ? W.u2 : u;
            if (prm->const_evals)
                hipLaunchKernelGGL(cmf_p_kernel<true>, dim3(nb), dim3(mh::kBlock), 0, s, hp, tmp, W.q4,
                                   eigs, alpha, D, PP, W.ctrl);
            else
                hipLaunchKernelGGL(cmf_p_kernel<false>, dim3(nb), dim3(mh::kBlock), 0, s, hp, tmp, W.q4,
                                   eigs, alpha, D, PP, W.ctrl);
            const bool chk = (it % prm->update) == 0;
            if (sup && chk)
                hipLaunchKernelGGL((cmf_u_kernel<true, true>), dim3(nb), dim3(mh::kBlock), 0, s, hp, tmp,
                                   cs, ct, pt, uu, labels, D, UP, W.ctrl, W.partials, W.crit_log);
            else if (sup)
                hipLaunchKernelGGL((cmf_u_kernel<true, false>), dim3(nb), dim3(mh::kBlock), 0, s, hp, tmp,
                                   cs, ct, pt, uu, labels, D, UP, W.ctrl, W.partials, W.crit_log);
            else if (chk)
                hipLaunchKernelGGL((cmf_u_kernel<false, true>), dim3(nb), dim3(mh::kBlock), 0, s, hp, tmp,
                                   cs, ct, pt, uu, labels, D, UP, W.ctrl, W.partials, W.crit_log);
            else
                hipLaunchKernelGGL((cmf_u_kernel<false, false>), dim3(nb), dim3(mh::kBlock), 0, s, hp,
                                   tmp, cs, ct, pt, uu, labels, D, UP, W.ctrl, W.partials, W.crit_log);
            if (chk)
                hipLaunchKernelGGL(cmf_crit_kernel, dim3(1), dim3(1024), 0, s, W.partials, int(nb),
                                   UP.inv_n_rcc, UP.error_bound, W.ctrl, W.crit_log);
            it += 1;
            return 1;
        }
        // a criterion iteration is always the last level of its pass, so a
        // stop never skips an iteration the pass already applied; passes of
        // fewer levels than the shape takes run shape 0 or kShape3
        const int lim = std::min(prm->num_iter, limit);
        const int lv = sweep_levels(kSweepShapes[m.shape].levels, it, lim, prm->update);
        const bool chk = ((it + lv - 1) % prm->update) == 0;
        SweepBufs SB;
        if (cur == 0) SB = SweepBufs{W.hp, W.tmp, W.pt, u, W.hp2, W.tmp2, W.pt2, W.u2};
        else SB = SweepBufs{W.hp2, W.tmp2, W.pt2, W.u2, W.hp, W.tmp, W.pt, u};
        const SweepMethod mm = pass_method(m, lv);
        const int64_t grid = sweep_grid_of(mm);
        launch_sweep(mm.shape, unsigned(grid), s, SB, W.q4, alpha, cs, ct, D, PP, UP, mm.zchunk, lv,
                     int(chk), 1 - cur, W.ctrl, W.partials);
        if (chk)
            hipLaunchKernelGGL(cmf_crit_kernel, dim3(1), dim3(1024), 0, s, W.partials, int(grid),
                               UP.inv_n_rcc, UP.error_bound, W.ctrl, W.crit_log);
        cur ^= 1;
        swept = true;
        it += lv;
        return lv;
    }

    // The method for this solve: forced by the context, tuned earlier for
    // these dims, tuned now on this solve's own iterations (at least 96
    // iterations; tune 2: always, tune 1: when these dims were solved on this
    // context before, so that a shape met once, as in a ragged set, does not
    // pay for timing every candidate), or the default (the three-level sweep
    // with the 27 x 27 tile shape when eligible and the volume has >= 800,000
    // voxels, else the plain kernels).
    int choose(int tune, SweepMethod &use)
    {
        use = SweepMethod{-1, 0};
        const std::array<int, 3> key{D.d0, D.d1, D.d2};
        if (ctx->sweep_mode == 1 || !sweep_elig) return MH_OK;
        if (ctx->sweep_mode == 2) {
            const int k = ctx->sweep_shape > 0 ? std::min(ctx->sweep_shape, kNumSweepShapes) - 1 : 0;
            SweepMethod m{k, ctx->sweep_zchunk > 0 ? ctx->sweep_zchunk : model_zchunk(k)};
            if (usable(m)) use = m;
            return MH_OK;
        }
        auto f = ctx->sweep_tuned.find(key);
        if (f != ctx->sweep_tuned.end()) {
            use = SweepMethod{f->second.first, f->second.second};
            return MH_OK;
        }
        bool allow_tune = tune == 2;
        if (tune == 1) allow_tune = !ctx->sweep_seen.insert(key).second;
        if (!allow_tune || prm->num_iter < 96) {
            // measured: the sweep wins from about 10^6 voxels (100^3) on, the
            // per-iteration kernels at 80^3; the 27 x 27 three-level shape
            // was the fastest single solve at 128^3 and within 5 % of the
            // fastest at 256^3
            const SweepMethod m{kDefaultShape, model_zchunk(kDefaultShape)};
            if (usable(m) && n >= 800000) use = m;
            return MH_OK;
        }
        // candidates: the per-iteration kernels and every tile shape with its
        // modelled plane chunk
        SweepMethod cands[1 + kNumSweepShapes];
        cands[0] = SweepMethod{-1, 0};
        for (int k = 0; k < kNumSweepShapes; ++k) cands[1 + k] = SweepMethod{k, model_zchunk(k)};
        // time each candidate on this solve's own iterations: one warm-up
        // pass, then >= kTuneIters iterations
        ScopedEvent se0, se1;
        MH_CHECK(ctx, hipEventCreate(&se0.e));
        MH_CHECK(ctx, hipEventCreate(&se1.e));
        hipEvent_t e0 = se0.e, e1 = se1.e;
        float best = 0.f;
        for (const SweepMethod &m : cands) {
            if (!usable(m) || it + kTuneIters + 4 > prm->num_iter) continue;
            step(m);
            MH_CHECK(ctx, hipEventRecord(e0, s));
            int done = 0;
            while (done < kTuneIters) done += step(m);
            MH_CHECK(ctx, hipEventRecord(e1, s));
            MH_CHECK(ctx, hipEventSynchronize(e1));
            float ms = 0.f;
            MH_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
            const float per = ms / float(done);
            if (best == 0.f || per < best) {
                best = per;
                use = m;
            }
        }
        MH_CHECK_LAUNCH(ctx, "cmf tuning");
        int stopped = 0;
        MH_CHECK(ctx, hipMemcpy(&stopped, W.ctrl + kStop, sizeof(int), hipMemcpyDeviceToHost));
        // timings of a stopped solve (empty launches) choose nothing
        if (!stopped) ctx->sweep_tuned[key] = {use.shape, use.zchunk};
        return MH_OK;
    }

    int run_until(const SweepMethod &m, int target)
    {
        limit = std::min(target, prm->num_iter);
        // With graphs on, the per-iteration kernels' criterion cycle (update
        // iterations from a criterion iteration: the same launches every
        // time) is captured once as a graph and replayed. The kernels read
        // the stop flag themselves, as launched ones do, so a stop inside a
        // replay ends the solve the same way. Off by default: on gfx950
        // back-to-back launches on one stream already run at the dispatch
        // rate, and replayed cycles measured 1-20% slower from 32^3 to 128^3.
        if (m.shape < 0 && ctx->graph_mode == 1 && graphs_allowed) {
            while (it < limit && it % prm->update != 0) step(m);
            const int cycles = (limit - it) / prm->update;
            if (cycles >= 2) {
                const int rc = replay_cycles(m, cycles);
                if (rc) return rc;
            }
        }
        while (it < limit) step(m);
        MH_CHECK_LAUNCH(ctx, "cmf kernels");
        return MH_OK;
    }

    // Capture one criterion cycle and launch it `cycles` times. Capture needs
    // a stream other than the legacy default one, so the work goes to the
    // context's own non-blocking stream, ordered after and before the
    // caller's stream by events. If capture is refused, nothing is lost: the
    // captured launches never ran and the loop launches them normally.
    int replay_cycles(const SweepMethod &m, int cycles)
    {
        if (!ctx->graph_stream)
            MH_CHECK(ctx, hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking));
        hipStream_t gs = ctx->graph_stream;
        ScopedEvent sb, sa;
        MH_CHECK(ctx, hipEventCreateWithFlags(&sb.e, hipEventDisableTiming));
        MH_CHECK(ctx, hipEventCreateWithFlags(&sa.e, hipEventDisableTiming));
        hipEvent_t before = sb.e, after = sa.e;
        MH_CHECK(ctx, hipEventRecord(before, s));
        MH_CHECK(ctx, hipStreamWaitEvent(gs, before, 0));
        const hipStream_t caller = s;
        s = gs;
        const int it0 = it;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        bool ok = hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            for (int k = 0; k < prm->update; ++k) step(m);
            ok = hipStreamEndCapture(gs, &graph) == hipSuccess && graph != nullptr;
            it = it0;
        }
        if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        hipError_t e = hipSuccess;
        for (int c = 0; ok && c < cycles && e == hipSuccess; ++c) {
            e = hipGraphLaunch(exec, gs);
            it += prm->update;
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();  // a refused capture leaves nothing behind
        s = caller;
        MH_CHECK(ctx, hipEventRecord(after, gs));
        MH_CHECK(ctx, hipStreamWaitEvent(s, after, 0));
        MH_CHECK(ctx, e);
        return MH_OK;
    }

    int finish()
    {
        if (swept) {
            hipLaunchKernelGGL(cmf_sweep_final_kernel, dim3(nb), dim3(mh::kBlock), 0, s, u, W.u2, n,
                               W.ctrl);
            MH_CHECK_LAUNCH(ctx, "cmf sweep final");
        }
        return MH_OK;
    }
};

// The iterations of one solve on workspace W: initialisation, the method
// chosen for these dims (tune as in Runner::choose: the default for a shape
// met the first time, tuned on the second), the final copy of u when a sweep
// left it in the second buffer set. Enqueues only, except while tuning.
int solve_core(mh_ctx *ctx, float *u, const float *eigs, const float *alpha, const float *cs,
               const float *ct, const uint8_t *labels, const Dims D, int64_t n,
               const mh_cmf_params *prm, const Workspace &W, hipStream_t s, int tune)
{
    Runner R(ctx, u, eigs, alpha, cs, ct, labels, D, n, prm, W, s);
    int rc = R.init();
    if (rc) return rc;
    SweepMethod use;
    rc = R.choose(tune, use);
    if (rc) return rc;
    rc = R.run_until(use, prm->num_iter);
    if (rc) return rc;
    return R.finish();
}

}  // namespace

namespace {
int read_result(mh_ctx *ctx, const Workspace &W, const mh_cmf_params *prm, mh_cmf_result *res,
                float *crit_host, hipStream_t s);
}  // namespace

MH_EXPORT int64_t mh_cmf_workspace_bytes(int64_t n)
{
    if (n <= 0 || n >= (int64_t(1) << 31)) return -1;
    return layout(n, nullptr, nullptr);
}

MH_EXPORT int mh_cmf_solve(mh_ctx *ctx, float *u, const float *eigs, const float *alpha,
                           const float *cs, const float *ct, const uint8_t *labels,
                           const int64_t dims[3], const mh_cmf_params *prm, void *work,
                           int64_t work_bytes, mh_cmf_result *res, float *crit_host, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, u && eigs && alpha && cs && ct && prm && work, "null pointer");
    MH_REQUIRE(ctx, prm->num_iter >= 0 && prm->update >= 1, "num_iter >= 0 and update >= 1 required");
    MH_REQUIRE(ctx, prm->rcc > 0.f, "rcc must be positive");
    const int64_t n = dims[0] * dims[1] * dims[2];
    MH_REQUIRE(ctx, work_bytes >= layout(n, nullptr, nullptr), "workspace too small");
    MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(work) & 255) == 0, "workspace must be 256-byte aligned");
    Workspace W;
    layout(n, static_cast<char *>(work), &W);
    hipStream_t s = (hipStream_t)stream;
    const Dims D{int(dims[0]), int(dims[1]), int(dims[2])};

    const int rc_core = solve_core(ctx, u, eigs, alpha, cs, ct, labels, D, n, prm, W, s, 1);
    if (rc_core) return rc_core;
    return read_result(ctx, W, prm, res, crit_host, s);
}

namespace {

// Blocks; the criterion log, stop flag and iteration count of a solve.
int read_result(mh_ctx *ctx, const Workspace &W, const mh_cmf_params *prm, mh_cmf_result *res,
                float *crit_host, hipStream_t s)
{
    int ctrl[kCtrlWords];
    MH_CHECK(ctx, hipMemcpyAsync(ctrl, W.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    const int ncrit = std::min(ctrl[kNcrit], kCritLogMax);
    float last = 0.f;
    if (ncrit > 0)
        MH_CHECK(ctx, hipMemcpy(&last, W.crit_log + ncrit - 1, sizeof(float), hipMemcpyDeviceToHost));
    if (crit_host && ncrit > 0)
        MH_CHECK(ctx, hipMemcpy(crit_host, W.crit_log, sizeof(float) * ncrit, hipMemcpyDeviceToHost));
    if (res) {
        res->stopped = ctrl[kStop];
        res->ncrit = ncrit;
        res->iterations = ctrl[kStop] ? (ctrl[kNcrit] - 1) * prm->update + 1 : prm->num_iter;
        res->last_crit = last;
    }
    return ctrl[kStop] == 1 ? MH_STOPPED_NAN : MH_OK;
}

Runner runner_of(mh_ctx *ctx, mh_cmf_resident *r, Workspace &W, hipStream_t s)
{
    const int64_t n = r->dims[0] * r->dims[1] * r->dims[2];
    layout(n, static_cast<char *>(r->work_dev), &W);
    Runner R(ctx, r->u_dev, r->eigs_dev, r->alpha_dev, r->cs_dev, r->ct_dev, nullptr,
             Dims{int(r->dims[0]), int(r->dims[1]), int(r->dims[2])}, n, &r->params, W, s);
    R.cur = r->cur;
    R.it = r->iter;
    R.swept = r->swept != 0;
    return R;
}

int resident_check(mh_ctx *ctx, const mh_cmf_resident *r)
{
    MH_REQUIRE(ctx, r && mh::dims_ok(r->dims), "bad resident solve");
    MH_REQUIRE(ctx, r->u_dev && r->eigs_dev && r->alpha_dev && r->cs_dev && r->ct_dev && r->work_dev,
               "null pointer");
    MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(r->work_dev) & 255) == 0,
               "workspace must be 256-byte aligned");
    const mh_cmf_params &P = r->params;
    MH_REQUIRE(ctx, P.num_iter >= 0 && P.update >= 1 && P.rcc > 0.f, "bad params");
    return MH_OK;
}

}  // namespace

MH_EXPORT int mh_cmf_resident_init(mh_ctx *ctx, mh_cmf_resident *r, void *stream)
{
    int rc = resident_check(ctx, r);
    if (rc) return rc;
    r->iter = 0;
    r->cur = 0;
    r->swept = 0;
    r->method = -2;
    r->zchunk = 0;
    Workspace W;
    Runner R = runner_of(ctx, r, W, (hipStream_t)stream);
    return R.init();
}

MH_EXPORT int mh_cmf_resident_step(mh_ctx *ctx, mh_cmf_resident *r, int iters, int allow_tune,
                                   void *stream)
{
    int rc = resident_check(ctx, r);
    if (rc) return rc;
    MH_REQUIRE(ctx, iters >= 0, "iters >= 0");
    Workspace W;
    Runner R = runner_of(ctx, r, W, (hipStream_t)stream);
    SweepMethod m{r->method, r->zchunk};
    if (r->method == -2) {
        rc = R.choose(allow_tune != 0 ? 2 : 0, m);
        if (rc) return rc;
        r->method = m.shape;
        r->zchunk = m.zchunk;
    }
    rc = R.run_until(m, r->iter + iters > R.it ? std::max(R.it, r->iter + iters) : R.it);
    r->iter = R.it;
    r->cur = R.cur;
    r->swept = R.swept ? 1 : 0;
    return rc;
}

MH_EXPORT int mh_cmf_resident_status(mh_ctx *ctx, const mh_cmf_resident *r, int *stopped,
                                     void *stream)
{
    int rc = resident_check(ctx, r);
    if (rc) return rc;
    MH_REQUIRE(ctx, stopped, "null pointer");
    Workspace W;
    const int64_t n = r->dims[0] * r->dims[1] * r->dims[2];
    layout(n, static_cast<char *>(r->work_dev), &W);
    hipStream_t s = (hipStream_t)stream;
    MH_CHECK(ctx, hipMemcpyAsync(stopped, W.ctrl + kStop, sizeof(int), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    return MH_OK;
}

MH_EXPORT int mh_cmf_resident_finish(mh_ctx *ctx, mh_cmf_resident *r, mh_cmf_result *res,
                                     float *crit_host, void *stream)
{
    int rc = resident_check(ctx, r);
    if (rc) return rc;
    Workspace W;
    Runner R = runner_of(ctx, r, W, (hipStream_t)stream);
    rc = R.finish();
    if (rc) return rc;
    return read_result(ctx, W, &r->params, res, crit_host, (hipStream_t)stream);
}

// =========================================================================
// mh_cmf_solve for callers that cannot allocate device memory. The solve's
// workspace (control block, criterion log, partial sums, both state sets of
// the multi-level sweep and the expanded frames: mh_cmf_workspace_bytes(n),
// about 60 B per voxel) lives in `scratch`, one file-scope device array per
// GPU. Its device address is looked up once per device when a context is
// created on it (mh_noalloc_setup, called by mh_create with the device
// current) and only that address is handed to kernels. Every call on a GPU
// shares the array: the control block is cleared by a kernel at the start of
// each solve and every other word is written before it is read, so a call
// never sees the previous call's values; calls that overlap on one GPU (two
// streams, two threads) race on it.
// =========================================================================

constexpr int64_t kNoAllocMaxVoxels = MH_CMF_NOALLOC_MAX_VOXELS;
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace

// Called by mh_create on ctx->device (current): resolves this device's copy
// of `scratch` once and records it in the context.
int mh_noalloc_setup(mh_ctx *ctx)
{
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return mh::set_error(ctx, MH_ERR_ARG, "mh_create", "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return mh::set_error(ctx, MH_ERR_HIP, "hipGetSymbolAddress(scratch)",
                                 e != hipSuccess ? hipGetErrorString(e) : "null address");
        g_noalloc_ptr[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_noalloc_ptr[ctx->device];
    return MH_OK;
}

MH_EXPORT int mh_cmf_solve_noalloc(mh_ctx *ctx, float *u, const float *eigs, const float *alpha,
                                   const float *cs, const float *ct, const uint8_t *labels,
                                   const int64_t dims[3], const mh_cmf_params *prm,
                                   mh_cmf_result *res, float *crit_host, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, u && eigs && alpha && cs && ct && prm, "null pointer");
    MH_REQUIRE(ctx, prm->num_iter >= 0 && prm->update >= 1, "num_iter >= 0 and update >= 1 required");
    MH_REQUIRE(ctx, prm->rcc > 0.f, "rcc must be positive");
    const int64_t n = dims[0] * dims[1] * dims[2];
    MH_REQUIRE(ctx, n <= kNoAllocMaxVoxels,
               "volume exceeds the no-allocation limit of MH_CMF_NOALLOC_MAX_VOXELS (2^20) voxels");
    MH_REQUIRE(ctx, layout(n, nullptr, nullptr) <= int64_t(kNoAllocScratchBytes),
               "workspace exceeds the no-allocation scratch array");
    if (!ctx->noalloc_scratch)
        return mh::set_error(ctx, MH_ERR_HIP, __func__,
                             ctx->noalloc_error.empty() ? "scratch array not set up"
                                                        : ctx->noalloc_error.c_str());
    int cur_dev = -1;
    MH_CHECK(ctx, hipGetDevice(&cur_dev));
    MH_REQUIRE(ctx, cur_dev == ctx->device, "the context's device must be current");
    Workspace W;
    layout(n, static_cast<char *>(ctx->noalloc_scratch), &W);
    hipStream_t s = (hipStream_t)stream;
    const Dims D{int(dims[0]), int(dims[1]), int(dims[2])};
    // the method mh_cmf_solve tuned for these dims on this context, else the
    // untuned default; never tuned here (tuning creates events) and never
    // captured as a graph
    Runner R(ctx, u, eigs, alpha, cs, ct, labels, D, n, prm, W, s);
    R.graphs_allowed = false;
    int rc = R.init();
    if (rc) return rc;
    SweepMethod use;
    rc = R.choose(0, use);
    if (rc) return rc;
    rc = R.run_until(use, prm->num_iter);
    if (rc) return rc;
    rc = R.finish();
    if (rc) return rc;
    return read_result(ctx, W, prm, res, crit_host, s);
}

// =========================================================================
// Batched solves: S unsupervised constant-eigenvalue solves on one volume
// that share the frame field, cs, ct and |grad I|^2. A thread owns one voxel
// and a group of SI solves: the shared fields and the rotation are loaded
// once per group, and each solve's alpha = w exp(-gms / decay^2) is formed
// in registers. Solve s keeps its own state (hp, pt, tmp, u) and its own
// stop flag and criterion log.
// =========================================================================
namespace {


struct BatchCtrl {
    int *stop;        // [S]
    int *ncrit;       // [S]
    unsigned *count;  // [groups]
    float *crit_log;  // [S][maxcheck]
    int maxcheck;
};



__global__ void __launch_bounds__(mh::kBlock)
cmf_batch_init_kernel(float *__restrict__ hp, float *__restrict__ pt, float *__restrict__ tmp,
                      const float *__restrict__ u, const float *__restrict__ cs,
                      const float *__restrict__ ct, int64_t n, float rcc)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t s = blockIdx.y;
    float *h = hp + 3 * n * s;
    h[i] = 0.f;
    h[n + i] = 0.f;
    h[2 * n + i] = 0.f;
    pt[n * s + i] = fminf(cs[i], ct[i]);
    tmp[n * s + i] = -(u[n * s + i] * rcc);
}

__global__ void __launch_bounds__(mh::kBlock)
quat_kernel(float4 *__restrict__ q4, const float *__restrict__ eigs, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const mh::Quat q = mh::expand_squeezed(eigs[3 * i], eigs[3 * i + 1], eigs[3 * i + 2]);
    q4[i] = make_float4(q.w, q.x, q.y, q.z);
}

// NS solves per thread. Every load of all NS solves is issued before any
// arithmetic (the addresses of absent neighbours fall back to the voxel
// itself, and the selects below restore the boundary rules), so a thread
// keeps 7 NS loads in flight; a stopped solve is computed but not stored.
template <int NS>
__global__ void __launch_bounds__(mh::kBlock)
cmf_batch_p_kernel(float *__restrict__ hp, const float *__restrict__ tmp,
                   const float4 *__restrict__ q4, const float *__restrict__ gms, const Dims D,
                   const PParams P, const SolveParams SP, int s0, const int *__restrict__ stop)
{
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int i0, i1, i2;
    coords(i, D, i0, i1, i2);
    const int64_t st0 = int64_t(D.d1) * D.d2;
    const bool e0 = i0 < D.d0 - 1, e1 = i1 < D.d1 - 1, e2 = i2 < D.d2 - 1;
    const int64_t j0 = e0 ? i + st0 : i, j1 = e1 ? i + D.d2 : i, j2 = e2 ? i + 1 : i;
    const float4 qq = q4[i];
    const float gm = gms[i];
    float t[NS], a0[NS], a1[NS], a2[NS], h0[NS], h1[NS], h2[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const float *ts = tmp + n * (s0 + k);
        const float *h = hp + 3 * n * (s0 + k);
        t[k] = ts[i];
        a0[k] = ts[j0];
        a1[k] = ts[j1];
        a2[k] = ts[j2];
        h0[k] = h[i];
        h1[k] = h[n + i];
        h2[k] = h[2 * n + i];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if (stop[s0 + k]) continue;  // uniform over the grid
        const float g0 = e0 ? (a0[k] - t[k]) * P.ss0 : 0.f;
        const float g1 = e1 ? (a1[k] - t[k]) * P.ss1 : 0.f;
        const float g2 = e2 ? (a2[k] - t[k]) * P.ss2 : 0.f;
        mh::cmf_p_voxel(h0[k], h1[k], h2[k], g0, g1, g2, qq, P.sq0, P.sq1, P.sq2, P.rq0, P.rq1, P.rq2,
                        P.A, SP.w[k] * expf(-gm * SP.k[k]));
        float *h = hp + 3 * n * (s0 + k);
        h[i] = h0[k];
        h[n + i] = h1[k];
        h[2 * n + i] = h2[k];
    }
}

template <int NS, bool CHECK>
__global__ void __launch_bounds__(mh::kBlock)
cmf_batch_u_kernel(const float *__restrict__ hp, float *__restrict__ tmp,
                   const float *__restrict__ cs, const float *__restrict__ ct,
                   float *__restrict__ pt, float *__restrict__ u, const Dims D, const UParams P,
                   int s0, BatchCtrl C, float *__restrict__ partials)
{
    const int64_t n = int64_t(D.d0) * D.d1 * D.d2;
    const unsigned b = mh::xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = int64_t(b) * blockDim.x + threadIdx.x;
    float aerr[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) aerr[k] = 0.f;
    if (i < n) {
        int i0, i1, i2;
        coords(i, D, i0, i1, i2);
        const int64_t st0 = int64_t(D.d1) * D.d2;
        const bool e0 = i0 < D.d0 - 1, e1 = i1 < D.d1 - 1, e2 = i2 < D.d2 - 1;
        const bool f0 = i0 > 0, f1 = i1 > 0, f2 = i2 > 0;
        const int64_t k0 = f0 ? i - st0 : i, k1 = f1 ? i - D.d2 : i, k2 = f2 ? i - 1 : i;
        const float csi = cs[i], cti = ct[i];
        float ha0[NS], ha1[NS], ha2[NS], hb0[NS], hb1[NS], hb2[NS], uu[NS], pp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const float *h = hp + 3 * n * (s0 + k);
            ha0[k] = h[i];
            hb0[k] = h[k0];
            ha1[k] = h[n + i];
            hb1[k] = h[n + k1];
            ha2[k] = h[2 * n + i];
            hb2[k] = h[2 * n + k2];
            uu[k] = u[n * (s0 + k) + i];
            pp[k] = pt[n * (s0 + k) + i];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if (C.stop[s0 + k]) continue;  // uniform over the grid
            const float x0 = (e0 ? ha0[k] : 0.f) - (f0 ? hb0[k] : 0.f);
            const float x1 = (e1 ? ha1[k] : 0.f) - (f1 ? hb1[k] : 0.f);
            const float x2 = (e2 ? ha2[k] : 0.f) - (f2 ? hb2[k] : 0.f);
            const float div = mh::cmf_div(x0, x1, x2, P.rs0, P.rs1, P.rs2);
            float ti;
            const float err = mh::cmf_u_voxel(div, pp[k], uu[k], ti, csi, cti, P.rcc, P.inv_rcc, 1.f, 0.f);
            const int64_t o = n * (s0 + k) + i;
            pt[o] = pp[k];
            u[o] = uu[k];
            tmp[o] = ti;
            aerr[k] = fabsf(err);
        }
    }
    if (!CHECK) return;
    const unsigned nbx = gridDim.x;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const float v = mh::block_reduce<SumOp>(aerr[k]);
        if (threadIdx.x == 0) partials[int64_t(s0 + k) * nbx + blockIdx.x] = v;
    }
}

// One work-group per solve: criterion from its partials, log, stop flag.
__global__ void __launch_bounds__(mh::kBlock)
cmf_batch_crit_kernel(const float *__restrict__ partials, int nparts, float inv_n_rcc,
                      float error_bound, BatchCtrl C, int s0)
{
    const int s = s0 + blockIdx.x;
    if (C.stop[s]) return;
    float acc = 0.f;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += partials[int64_t(s) * nparts + q];
    acc = mh::block_reduce<SumOp>(acc);
    if (threadIdx.x == 0) {
        const float crit = acc * inv_n_rcc;
        const int c = C.ncrit[s];
        if (c < C.maxcheck) C.crit_log[int64_t(s) * C.maxcheck + c] = crit;
        C.ncrit[s] = c + 1;
        if (crit != crit) C.stop[s] = 1;
        else if (crit < error_bound) C.stop[s] = 2;
    }
}

struct BatchWorkspace {
    int *stop, *ncrit, *parity;
    unsigned *count;
    float *crit_log, *partials, *hp, *pt, *tmp;
    float4 *q4;
    // second state set of one group for the multi-level sweep
    float *hp2, *pt2, *tmp2, *u2, *alpha2;  // and the group's alpha fields
    int maxcheck;
};

int64_t batch_layout(int64_t n, int S, int maxcheck, char *base, BatchWorkspace *w)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    const int groups = (S + kSI - 1) / kSI;
    BatchWorkspace ws;
    ws.stop = reinterpret_cast<int *>(take(sizeof(int) * (2 * int64_t(S) + groups)));
    ws.ncrit = ws.stop ? ws.stop + S : nullptr;
    ws.count = ws.stop ? reinterpret_cast<unsigned *>(ws.stop + 2 * S) : nullptr;
    ws.crit_log = reinterpret_cast<float *>(take(sizeof(float) * int64_t(S) * maxcheck));
    ws.partials = reinterpret_cast<float *>(take(sizeof(float) * int64_t(S) * mh::blocks_for(n)));
    ws.q4 = reinterpret_cast<float4 *>(take(sizeof(float4) * n));
    ws.hp = reinterpret_cast<float *>(take(sizeof(float) * 3 * n * S));
    ws.pt = reinterpret_cast<float *>(take(sizeof(float) * n * S));
    ws.tmp = reinterpret_cast<float *>(take(sizeof(float) * n * S));
    ws.parity = reinterpret_cast<int *>(take(sizeof(int) * int64_t(S)));
    const int64_t g = std::min(S, kSI);
    ws.hp2 = reinterpret_cast<float *>(take(sizeof(float) * 3 * n * g));
    ws.pt2 = reinterpret_cast<float *>(take(sizeof(float) * n * g));
    ws.tmp2 = reinterpret_cast<float *>(take(sizeof(float) * n * g));
    ws.u2 = reinterpret_cast<float *>(take(sizeof(float) * n * g));
    ws.alpha2 = reinterpret_cast<float *>(take(sizeof(float) * n * g));
    ws.maxcheck = maxcheck;
    if (w) *w = ws;
    return off;
}

template <int NS>
void launch_batch_ns(bool chk, unsigned nb, hipStream_t s, const BatchWorkspace &W,
                     const float *gms, const float *cs, const float *ct, float *u, const Dims &D,
                     const PParams &PP, const UParams &UP, const SolveParams &sp, int s0,
                     const BatchCtrl &C)
{
    hipLaunchKernelGGL(cmf_batch_p_kernel<NS>, dim3(nb), dim3(mh::kBlock), 0, s, W.hp, W.tmp, W.q4,
                       gms, D, PP, sp, s0, W.stop);
    if (chk)
        hipLaunchKernelGGL((cmf_batch_u_kernel<NS, true>), dim3(nb), dim3(mh::kBlock), 0, s, W.hp,
                           W.tmp, cs, ct, W.pt, u, D, UP, s0, C, W.partials);
    else
        hipLaunchKernelGGL((cmf_batch_u_kernel<NS, false>), dim3(nb), dim3(mh::kBlock), 0, s, W.hp,
                           W.tmp, cs, ct, W.pt, u, D, UP, s0, C, W.partials);
}

// p- and u-step of one group of ns (1..kSI) solves starting at solve s0.
void launch_batch_group(int ns, bool chk, unsigned nb, hipStream_t s, const BatchWorkspace &W,
                        const float *gms, const float *cs, const float *ct, float *u,
                        const Dims &D, const PParams &PP, const UParams &UP,
                        const SolveParams &sp, int s0, const BatchCtrl &C)
{
    switch (ns) {
    case 1: launch_batch_ns<1>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 2: launch_batch_ns<2>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 3: launch_batch_ns<3>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 4: launch_batch_ns<4>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 5: launch_batch_ns<5>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 6: launch_batch_ns<6>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    case 7: launch_batch_ns<7>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    default: launch_batch_ns<kSI>(chk, nb, s, W, gms, cs, ct, u, D, PP, UP, sp, s0, C); break;
    }
}

int maxcheck_for(const mh_cmf_params *p)
{
    return p->update >= 1 ? std::max(1, (p->num_iter + p->update - 1) / p->update) : 1;
}

}  // namespace

namespace {

// Tune the single-solve method for these dims on a probe: solve 0's fields
// (u0, alpha from |grad I|^2) copied into scratch memory and 96 iterations of
// a solve that cannot stop early; the context keeps the choice.
int tune_on_probe(mh_ctx *ctx, const float *u, float w, float decay, const float *eigs,
                  const float *gms, const float *cs, const float *ct, const Dims D, int64_t n,
                  const mh_cmf_params *prm, hipStream_t s)
{
    const int64_t wbytes = layout(n, nullptr, nullptr);
    char *mem = nullptr;
    MH_CHECK(ctx, hipMalloc(&mem, wbytes + 2 * align256(sizeof(float) * n)));
    float *us = reinterpret_cast<float *>(mem + wbytes);
    float *as = reinterpret_cast<float *>(mem + wbytes + align256(sizeof(float) * n));
    SolveParams sp{};
    sp.w[0] = w;
    sp.k[0] = float(1.0 / (double(decay) * double(decay)));
    int rc = mh::check_hip(ctx, hipMemcpyAsync(us, u, sizeof(float) * n, hipMemcpyDeviceToDevice, s),
                           "probe copy");
    if (!rc) {
        hipLaunchKernelGGL(cmf_group_alpha_kernel, dim3(mh::blocks_for(n), 1), dim3(mh::kBlock), 0, s,
                           gms, n, sp, as);
        Workspace W;
        layout(n, mem, &W);
        mh_cmf_params probe = *prm;
        probe.num_iter = 10 * (kTuneIters + 4) + 32;  // every candidate's window
        probe.error_bound = 0.f;
        rc = solve_core(ctx, us, eigs, as, cs, ct, nullptr, D, n, &probe, W, s, 2);
    }
    const hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(mem);
    if (!rc && e != hipSuccess) rc = mh::check_hip(ctx, e, "probe");
    return rc;
}

}  // namespace

MH_EXPORT int64_t mh_cmf_batched_workspace_bytes(int64_t n, int nsolves)
{
    if (n <= 0 || n >= (int64_t(1) << 31) || nsolves <= 0) return -1;
    // the criterion log is sized for the largest check count a call may use
    return batch_layout(n, nsolves, kCritLogMax / 16, nullptr, nullptr);
}

MH_EXPORT int mh_cmf_solve_batched(mh_ctx *ctx, float *u, int S, const float *w_host,
                                   const float *decay_host, const float *eigs, const float *gms,
                                   const float *cs, const float *ct, const int64_t dims[3],
                                   const mh_cmf_params *prm, void *work, int64_t work_bytes,
                                   mh_cmf_result *res, void *stream)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims), "bad dims");
    MH_REQUIRE(ctx, u && w_host && decay_host && eigs && gms && cs && ct && prm && work, "null pointer");
    MH_REQUIRE(ctx, S >= 1, "nsolves must be >= 1");
    MH_REQUIRE(ctx, prm->const_evals, "batched solves use constant eigenvalues");
    MH_REQUIRE(ctx, prm->num_iter >= 0 && prm->update >= 1 && prm->rcc > 0.f, "bad params");
    const int maxcheck = maxcheck_for(prm);
    MH_REQUIRE(ctx, maxcheck <= kCritLogMax / 16, "too many criterion checks (num_iter / update)");
    const int64_t n = dims[0] * dims[1] * dims[2];
    MH_REQUIRE(ctx, work_bytes >= batch_layout(n, S, kCritLogMax / 16, nullptr, nullptr), "workspace too small");
    MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(work) & 255) == 0, "workspace must be 256-byte aligned");
    BatchWorkspace W;
    batch_layout(n, S, maxcheck, static_cast<char *>(work), &W);
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = mh::blocks_for(n);
    const Dims D{int(dims[0]), int(dims[1]), int(dims[2])};
    const int groups = (S + kSI - 1) / kSI;

    const int nctrl = 2 * S + groups;
    hipLaunchKernelGGL(zero_ints, dim3((nctrl + 255) / 256), dim3(256), 0, s, W.stop, nctrl);
    hipLaunchKernelGGL(quat_kernel, dim3(nb), dim3(mh::kBlock), 0, s, W.q4, eigs, n);
    MH_CHECK_LAUNCH(ctx, "batched init");

    const PParams PP = p_params(prm);
    const UParams UP{prm->rcc, float(1.0 / double(prm->rcc)), prm->rs[0], prm->rs[1], prm->rs[2], prm->error_bound,
                     float(1.0 / (double(n) * double(prm->rcc)))};
    const BatchCtrl C{W.stop, W.ncrit, W.count, W.crit_log, maxcheck};
    // Schedule: groups of G solves (one thread handles a voxel of all G); a
    // group advances V iterations before the next group takes its turn. V =
    // num_iter runs the groups one after another, so a group's state stays in
    // the last-level cache between its iterations when it fits there.
    int G = ctx->batch_group, V = ctx->batch_visit;
    if (V <= 0) V = std::max(1, prm->num_iter);
    // Method. The multi-level sweep runs a group's iterations in passes over two
    // buffer sets, so it needs the groups one after another. Automatic: solve
    // one at a time through the sweep when the tuner of single solves chose a
    // sweep for these dims (tuning now on solve 0's fields in scratch memory
    // if it has not), else groups of 4 through the per-iteration kernels,
    // which share the frame and capacities in registers.
    const std::array<int, 3> key{D.d0, D.d1, D.d2};
    const bool sweep_elig = V >= prm->num_iter && n < kSweepMaxVoxels && ctx->batch_sweep != 1 &&
                            mh::axi_params(prm).on;
    bool want_sweep = sweep_elig && ctx->batch_sweep == 2;
    bool tune_pairs = false;
    SweepMethod sm{0, 16};
    if (ctx->sweep_mode == 2) {
        // the forced single-solve method, applied to every solve in turn
        // unless the batched sweep or the group kernels are asked for
        const int k = ctx->sweep_shape > 0 ? std::min(ctx->sweep_shape, kNumSweepShapes) - 1 : 0;
        sm = SweepMethod{k, ctx->sweep_zchunk > 0 ? ctx->sweep_zchunk
                                                   : sweep_model_zchunk(ctx, k, D, int64_t(nb))};
        if (sweep_elig && ctx->batch_sweep == 0) {
            want_sweep = true;
            if (G <= 0) G = sweep_pairs(sm.shape, n) ? 2 : 1;
        }
    } else {
        if (sweep_elig && ctx->batch_sweep == 0 && ctx->sweep_mode == 0 && prm->num_iter >= 96 &&
            ctx->sweep_tuned.find(key) == ctx->sweep_tuned.end()) {
            const int rc = tune_on_probe(ctx, u, w_host[0], decay_host[0], eigs, gms, cs, ct, D, n, prm, s);
            if (rc) return rc;
        }
        auto f = ctx->sweep_tuned.find(key);
        if (f != ctx->sweep_tuned.end() && f->second.first >= 0) {
            sm = SweepMethod{f->second.first, f->second.second};
            if (sweep_elig && ctx->batch_sweep == 0 && ctx->sweep_mode == 0) {
                want_sweep = true;
                // pairs run at other relative speeds than single solves: the
                // pair-capable shapes are timed on the first group's pair
                // (below) unless that was done for these dims
                auto fb2 = ctx->batch_tuned.find(key);
                if (n < kSweepPairMaxVoxels && S >= 2) {
                    if (fb2 != ctx->batch_tuned.end()) sm = SweepMethod{fb2->second.first, fb2->second.second};
                    else tune_pairs = prm->num_iter >= 96;
                }
                if (G <= 0) G = (sweep_pairs(sm.shape, n) || tune_pairs) ? 2 : 1;
            }
        }
    }
    if (G <= 0) G = 4;
    G = std::min(G, kSI);
    const int ngroups = (S + G - 1) / G;
    std::vector<SolveParams> gp(ngroups);
    for (int g = 0; g < ngroups; ++g)
        for (int k = 0; k < kSI; ++k) {
            const int j = std::min(g * G + k, S - 1);
            gp[g].w[k] = w_host[j];
            gp[g].k[k] = float(1.0 / (double(decay_host[j]) * double(decay_host[j])));
        }
    hipLaunchKernelGGL(zero_ints, dim3((S + 255) / 256), dim3(256), 0, s, W.parity, S);
    const SweepShape &SS = kSweepShapes[sm.shape];
    const int64_t nper = int64_t((D.d2 + SS.tx - 1) / SS.tx) * ((D.d1 + SS.ty - 1) / SS.ty) *
                         ((D.d0 + sm.zchunk - 1) / sm.zchunk);
    auto groups_of = [&](const SweepMethod &m) {
        const SweepShape &X = kSweepShapes[m.shape];
        return int64_t((D.d2 + X.tx - 1) / X.tx) * ((D.d1 + X.ty - 1) / X.ty) *
               ((D.d0 + m.zchunk - 1) / m.zchunk);
    };
    const SweepMethod fb{0, sweep_model_zchunk(ctx, 0, D, int64_t(nb))};
    const SweepMethod fb3{kShape3, sweep_model_zchunk(ctx, kShape3, D, int64_t(nb))};
    const int64_t nper_fb = groups_of(fb), nper_fb3 = groups_of(fb3);
    const bool use_sweep = want_sweep && nper <= int64_t(nb) && nper_fb <= int64_t(nb) &&
                           nper_fb3 <= int64_t(nb);
    // Pair tuning: the first group's passes try the tuned single-solve method
    // and each pair-capable shape (one warm-up pass, then >= 15 iterations
    // timed with events); the fastest per iteration runs the rest. Results do
    // not depend on the methods, bit for bit.
    std::vector<SweepMethod> cands;
    if (use_sweep && tune_pairs) {
        // the single-solve choice first, then the three-level shapes (the
        // fastest pairs measured), so that a solve too short to time every
        // candidate still times the likely ones
        cands.push_back(sm);
        for (int lv = 3; lv >= 2; --lv)
            for (int k = 0; k < kNumSweepShapes; ++k) {
                if (!sweep_pairs(k, n) || kSweepShapes[k].levels != lv) continue;
                const SweepMethod m{k, sweep_model_zchunk(ctx, k, D, int64_t(nb))};
                if (k != sm.shape && groups_of(m) <= int64_t(nb)) cands.push_back(m);
            }
    }
    size_t ci = 0;
    int tune_phase = 0, tune_it0 = 0;
    float tune_best = 0.f;
    SweepMethod tune_pick = sm;
    ScopedEvent ste0, ste1;
    if (!cands.empty()) {
        MH_CHECK(ctx, hipEventCreate(&ste0.e));
        MH_CHECK(ctx, hipEventCreate(&ste1.e));
    }
    hipEvent_t te0 = ste0.e, te1 = ste1.e;
    if (!use_sweep) {
        hipLaunchKernelGGL(cmf_batch_init_kernel, dim3(nb, S), dim3(mh::kBlock), 0, s, W.hp, W.pt,
                           W.tmp, u, cs, ct, n, prm->rcc);
        MH_CHECK_LAUNCH(ctx, "batched init");
    }
    // The sweep runs the groups one after another, so only the running
    // group's state is live: it is set up just before the group runs, in one
    // block at the start of the state arrays when they have room for four
    // groups (the state of all S solves would put one group's buffers
    // gigabytes apart: 5 % slower on the 425-solve grid at 128^3).
    const bool compact = int64_t(S) >= 4 * int64_t(G);
    for (int g = 0; use_sweep && g < ngroups; ++g) {
        const int s0 = g * G, ns = std::min(G, S - s0);
        const int64_t o = n * s0, nG = n * G;
        float *hpA = W.hp + 3 * o, *ptA = W.pt + o, *tmpA = W.tmp + o;
        float *hpB = W.hp2, *ptB = W.pt2, *tmpB = W.tmp2, *uB = W.u2, *alB = W.alpha2;
        if (compact) {  // 12 n G <= 3 n S floats of the hp array
            hpA = W.hp;
            ptA = hpA + 3 * nG;
            tmpA = ptA + nG;
            hpB = tmpA + nG;
            ptB = hpB + 3 * nG;
            tmpB = ptB + nG;
            uB = tmpB + nG;
            alB = uB + nG;
        }
        hipLaunchKernelGGL(cmf_batch_init_kernel, dim3(nb, ns), dim3(mh::kBlock), 0, s, hpA, ptA, tmpA,
                           u + o, cs, ct, n, prm->rcc);
        hipLaunchKernelGGL(cmf_group_alpha_kernel, dim3(nb, ns), dim3(mh::kBlock), 0, s, gms, n, gp[g],
                           alB);
        int cur = 0;
        for (int it = 0; it < prm->num_iter;) {
            const bool tuning = g == 0 && ci < cands.size();
            const SweepMethod mth = tuning ? cands[ci] : sm;
            const int levels = kSweepShapes[mth.shape].levels;
            const int lv = sweep_levels(levels, it, prm->num_iter, prm->update);
            const bool chk = ((it + lv - 1) % prm->update) == 0;
            // a shorter pass than the method's levels runs shape 0 or kShape3
            const bool own = lv == levels || (levels == 2 && lv < 2);
            const SweepMethod pm = own ? mth : (lv <= 2 ? fb : fb3);
            const int64_t pn = groups_of(pm);
            SweepBufs B0{hpA, tmpA, ptA, u + o, hpB, tmpB, ptB, uB};
            SweepBufs B1{hpB, tmpB, ptB, uB, hpA, tmpA, ptA, u + o};
            // the group's solves two per thread, an odd last one alone
            const int npair = sweep_pairs(pm.shape, n) ? ns / 2 : 0, rest = ns - 2 * npair;
            if (npair) {
                const SweepBatch SBp{s0, 2 * npair, int(pn), W.stop, W.parity};
                launch_sweep(pm.shape, unsigned(pn * npair), s, cur == 0 ? B0 : B1, W.q4, alB, cs, ct, D,
                             PP, UP, pm.zchunk, lv, int(chk), 1 - cur, nullptr, W.partials, &SBp, 2);
            }
            if (rest) {
                const int64_t r = n * (2 * npair);
                SweepBufs Br = cur == 0 ? B0 : B1;
                Br.hp += 3 * r;
                Br.tmp += r;
                Br.pt += r;
                Br.u += r;
                Br.hpo += 3 * r;
                Br.tmpo += r;
                Br.pto += r;
                Br.uo += r;
                const SweepBatch SBr{s0 + 2 * npair, rest, int(pn), W.stop, W.parity};
                launch_sweep(pm.shape, unsigned(pn * rest), s, Br, W.q4, alB + r, cs, ct, D, PP, UP,
                             pm.zchunk, lv, int(chk), 1 - cur, nullptr, W.partials, &SBr, 1);
            }
            if (chk)
                hipLaunchKernelGGL(cmf_batch_crit_kernel, dim3(ns), dim3(mh::kBlock), 0, s, W.partials,
                                   int(pn), UP.inv_n_rcc, UP.error_bound, C, s0);
            MH_CHECK_LAUNCH(ctx, "batched sweep");
            cur ^= 1;
            it += lv;
            if (tuning) {
                if (tune_phase == 0) {  // the warm-up pass is done
                    MH_CHECK(ctx, hipEventRecord(te0, s));
                    tune_phase = 1;
                    tune_it0 = it;
                } else if (it - tune_it0 >= kTuneIters) {
                    MH_CHECK(ctx, hipEventRecord(te1, s));
                    MH_CHECK(ctx, hipEventSynchronize(te1));
                    float ms = 0.f;
                    MH_CHECK(ctx, hipEventElapsedTime(&ms, te0, te1));
                    const float per = ms / float(it - tune_it0);
                    if (tune_best == 0.f || per < tune_best) {
                        tune_best = per;
                        tune_pick = mth;
                    }
                    tune_phase = 0;
                    if (++ci == cands.size()) {
                        sm = tune_pick;
                        // timings of a stopped solve (empty launches) choose nothing
                        std::vector<int> stopped(static_cast<size_t>(ns), 0);
                        MH_CHECK(ctx, hipMemcpy(stopped.data(), W.stop, sizeof(int) * size_t(ns),
                                                hipMemcpyDeviceToHost));
                        if (std::all_of(stopped.begin(), stopped.end(), [](int v) { return v == 0; }))
                            ctx->batch_tuned[key] = {sm.shape, sm.zchunk};
                    }
                }
            }
        }
        hipLaunchKernelGGL(cmf_batch_sweep_final_kernel, dim3(nb, ns), dim3(mh::kBlock), 0, s, u, uB, n,
                           s0, W.parity);
        MH_CHECK_LAUNCH(ctx, "batched sweep final");
        if (g == 0 && ci < cands.size()) sm = tune_pick;  // too few iterations to time them all
    }
    for (int base = 0; !use_sweep && base < prm->num_iter; base += V) {
        const int last = std::min(prm->num_iter, base + V);
        for (int g = 0; g < ngroups; ++g) {
            const int s0 = g * G, ns = std::min(G, S - s0);
            for (int it = base; it < last; ++it) {
                const bool chk = (it % prm->update) == 0;
                launch_batch_group(ns, chk, nb, s, W, gms, cs, ct, u, D, PP, UP, gp[g], s0, C);
                if (chk)
                    hipLaunchKernelGGL(cmf_batch_crit_kernel, dim3(ns), dim3(mh::kBlock), 0, s,
                                       W.partials, int(nb), UP.inv_n_rcc, UP.error_bound, C, s0);
            }
        }
        MH_CHECK_LAUNCH(ctx, "batched cmf kernels");
    }
    std::vector<int> stop(2 * S);
    MH_CHECK(ctx, hipMemcpyAsync(stop.data(), W.stop, sizeof(int) * 2 * S, hipMemcpyDeviceToHost, s));
    std::vector<float> log(int64_t(S) * maxcheck);
    MH_CHECK(ctx, hipMemcpyAsync(log.data(), W.crit_log, sizeof(float) * log.size(), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    int any_nan = 0;
    for (int j = 0; j < S; ++j) {
        const int st = stop[j], nc = std::min(stop[S + j], maxcheck);
        any_nan |= st == 1;
        if (res) {
            res[j].stopped = st;
            res[j].ncrit = nc;
            res[j].iterations = st ? (stop[S + j] - 1) * prm->update + 1 : prm->num_iter;
            res[j].last_crit = nc > 0 ? log[int64_t(j) * maxcheck + nc - 1] : 0.f;
        }
    }
    return any_nan ? MH_STOPPED_NAN : MH_OK;
}

// =========================================================================
// Ragged batch: V volumes of different dims packed one after another; every
// field of volume v starts at element offset off[v]. The grid covers the
// blocks of all volumes; block -> volume by binary search over the block
// prefix table. One unsupervised constant-eigenvalue solve per volume.
// =========================================================================
namespace {

struct Vol {
    int d0, d1, d2, block0;
    int64_t off;
};

__device__ __forceinline__ int find_vol(const Vol *__restrict__ vols, int nvol, unsigned b)
{
    int lo = 0, hi = nvol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (unsigned(vols[mid].block0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(mh::kBlock)
ragged_init_kernel(const Vol *__restrict__ vols, int nvol, float4 *__restrict__ q4,
                   const float *__restrict__ eigs, float *__restrict__ hp, float *__restrict__ pt,
                   float *__restrict__ tmp, const float *__restrict__ u, const float *__restrict__ cs,
                   const float *__restrict__ ct, float rcc)
{
    const int v = find_vol(vols, nvol, blockIdx.x);
    const Vol V = vols[v];
    const int64_t n = int64_t(V.d0) * V.d1 * V.d2;
    const int64_t li = int64_t(blockIdx.x - V.block0) * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const int64_t i = V.off + li;
    const mh::Quat q = mh::expand_squeezed(eigs[3 * i], eigs[3 * i + 1], eigs[3 * i + 2]);
    q4[i] = make_float4(q.w, q.x, q.y, q.z);
    float *h = hp + 3 * V.off;
    h[li] = 0.f;
    h[n + li] = 0.f;
    h[2 * n + li] = 0.f;
    pt[i] = fminf(cs[i], ct[i]);
    tmp[i] = -(u[i] * rcc);
}

__global__ void __launch_bounds__(mh::kBlock)
ragged_p_kernel(const Vol *__restrict__ vols, int nvol, float *__restrict__ hp,
                const float *__restrict__ tmp, const float4 *__restrict__ q4,
                const float *__restrict__ alpha, const PParams P, const int *__restrict__ stop)
{
    const int v = find_vol(vols, nvol, blockIdx.x);
    if (stop[v]) return;
    const Vol V = vols[v];
    const int64_t n = int64_t(V.d0) * V.d1 * V.d2;
    const int64_t li = int64_t(blockIdx.x - V.block0) * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const Dims D{V.d0, V.d1, V.d2};
    int i0, i1, i2;
    coords(li, D, i0, i1, i2);
    const int64_t i = V.off + li;
    const int64_t st0 = int64_t(V.d1) * V.d2;
    const float t = tmp[i];
    const float g0 = i0 < V.d0 - 1 ? (tmp[i + st0] - t) * P.ss0 : 0.f;
    const float g1 = i1 < V.d1 - 1 ? (tmp[i + V.d2] - t) * P.ss1 : 0.f;
    const float g2 = i2 < V.d2 - 1 ? (tmp[i + 1] - t) * P.ss2 : 0.f;
    float *h = hp + 3 * V.off;
    float h0 = h[li], h1 = h[n + li], h2 = h[2 * n + li];
    mh::cmf_p_voxel(h0, h1, h2, g0, g1, g2, q4[i], P.sq0, P.sq1, P.sq2, P.rq0, P.rq1, P.rq2, P.A,
                    alpha[i]);
    h[li] = h0;
    h[n + li] = h1;
    h[2 * n + li] = h2;
}

template <bool CHECK>
__global__ void __launch_bounds__(mh::kBlock)
ragged_u_kernel(const Vol *__restrict__ vols, int nvol, const float *__restrict__ hp,
                float *__restrict__ tmp, const float *__restrict__ cs, const float *__restrict__ ct,
                float *__restrict__ pt, float *__restrict__ u, const UParams P, BatchCtrl C,
                float *__restrict__ partials)
{
    const int v = find_vol(vols, nvol, blockIdx.x);
    if (C.stop[v]) return;  // uniform over the block
    const Vol V = vols[v];
    const int64_t n = int64_t(V.d0) * V.d1 * V.d2;
    const int64_t li = int64_t(blockIdx.x - V.block0) * blockDim.x + threadIdx.x;
    float aerr = 0.f;
    if (li < n) {
        const Dims D{V.d0, V.d1, V.d2};
        int i0, i1, i2;
        coords(li, D, i0, i1, i2);
        const int64_t i = V.off + li;
        const int64_t st0 = int64_t(V.d1) * V.d2;
        const float *h = hp + 3 * V.off;
        const float x0 = (i0 < V.d0 - 1 ? h[li] : 0.f) - (i0 > 0 ? h[li - st0] : 0.f);
        const float x1 = (i1 < V.d1 - 1 ? h[n + li] : 0.f) - (i1 > 0 ? h[n + li - V.d2] : 0.f);
        const float x2 = (i2 < V.d2 - 1 ? h[2 * n + li] : 0.f) - (i2 > 0 ? h[2 * n + li - 1] : 0.f);
        const float div = mh::cmf_div(x0, x1, x2, P.rs0, P.rs1, P.rs2);
        float ui = u[i], pti = pt[i], ti;
        const float err = mh::cmf_u_voxel(div, pti, ui, ti, cs[i], ct[i], P.rcc, P.inv_rcc, 1.f, 0.f);
        pt[i] = pti;
        u[i] = ui;
        tmp[i] = ti;
        aerr = fabsf(err);
    }
    if (!CHECK) return;
    aerr = mh::block_reduce<SumOp>(aerr);
    if (threadIdx.x == 0) partials[blockIdx.x] = aerr;
}

// One work-group per volume: criterion from the partials of its blocks.
__global__ void __launch_bounds__(mh::kBlock)
ragged_crit_kernel(const Vol *__restrict__ vols, const float *__restrict__ partials,
                   float rcc, float error_bound, BatchCtrl C)
{
    const int v = blockIdx.x;
    if (C.stop[v]) return;
    const Vol V = vols[v];
    const int64_t n = int64_t(V.d0) * V.d1 * V.d2;
    const int nbv = int((n + mh::kBlock - 1) / mh::kBlock);
    float acc = 0.f;
    for (int q = threadIdx.x; q < nbv; q += blockDim.x) acc += partials[V.block0 + q];
    acc = mh::block_reduce<SumOp>(acc);
    if (threadIdx.x == 0) {
        const float crit = acc / (float(n) * rcc);
        const int c = C.ncrit[v];
        if (c < C.maxcheck) C.crit_log[int64_t(v) * C.maxcheck + c] = crit;
        C.ncrit[v] = c + 1;
        if (crit != crit) C.stop[v] = 1;
        else if (crit < error_bound) C.stop[v] = 2;
    }
}

struct RaggedWorkspace {
    Vol *vols;
    int *stop, *ncrit;
    unsigned *count;
    float *crit_log, *partials, *hp, *pt, *tmp;
    float4 *q4;
};

int64_t ragged_layout(int64_t total, int nvol, int64_t nblocks, int maxcheck, char *base,
                      RaggedWorkspace *w)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    RaggedWorkspace ws;
    ws.vols = reinterpret_cast<Vol *>(take(sizeof(Vol) * nvol));
    ws.stop = reinterpret_cast<int *>(take(sizeof(int) * 3 * int64_t(nvol)));
    ws.ncrit = ws.stop ? ws.stop + nvol : nullptr;
    ws.count = ws.stop ? reinterpret_cast<unsigned *>(ws.stop + 2 * nvol) : nullptr;
    ws.crit_log = reinterpret_cast<float *>(take(sizeof(float) * int64_t(nvol) * maxcheck));
    ws.partials = reinterpret_cast<float *>(take(sizeof(float) * nblocks));
    ws.q4 = reinterpret_cast<float4 *>(take(sizeof(float4) * total));
    ws.hp = reinterpret_cast<float *>(take(sizeof(float) * 3 * total));
    ws.pt = reinterpret_cast<float *>(take(sizeof(float) * total));
    ws.tmp = reinterpret_cast<float *>(take(sizeof(float) * total));
    if (w) *w = ws;
    return off;
}

}  // namespace

MH_EXPORT int64_t mh_cmf_ragged_workspace_bytes(int64_t total, int nvol)
{
    if (total <= 0 || nvol <= 0) return -1;
    // every volume can add at most one partially filled block
    const int64_t nblocks = mh::blocks_for(total) + nvol;
    return ragged_layout(total, nvol, nblocks, kCritLogMax / 16, nullptr, nullptr);
}

MH_EXPORT int mh_cmf_solve_ragged(mh_ctx *ctx, float *u, const float *eigs, const float *alpha,
                                  const float *cs, const float *ct, int nvol, const int64_t *vdims,
                                  const int64_t *voff, const mh_cmf_params *prm, void *work,
                                  int64_t work_bytes, mh_cmf_result *res, void *stream)
{
    MH_REQUIRE(ctx, u && eigs && alpha && cs && ct && vdims && voff && prm && work, "null pointer");
    MH_REQUIRE(ctx, nvol >= 1, "nvol must be >= 1");
    MH_REQUIRE(ctx, prm->const_evals, "ragged solves use constant eigenvalues");
    MH_REQUIRE(ctx, prm->num_iter >= 0 && prm->update >= 1 && prm->rcc > 0.f, "bad params");
    const int maxcheck = maxcheck_for(prm);
    MH_REQUIRE(ctx, maxcheck <= kCritLogMax / 16, "too many criterion checks (num_iter / update)");
    std::vector<Vol> vols(nvol);
    int64_t total = 0, nblocks = 0;
    for (int v = 0; v < nvol; ++v) {
        MH_REQUIRE(ctx, mh::dims_ok(vdims + 3 * v), "bad volume dims");
        const int64_t n = vdims[3 * v] * vdims[3 * v + 1] * vdims[3 * v + 2];
        MH_REQUIRE(ctx, voff[v] >= 0, "negative volume offset");
        vols[v] = Vol{int(vdims[3 * v]), int(vdims[3 * v + 1]), int(vdims[3 * v + 2]), int(nblocks), voff[v]};
        nblocks += mh::blocks_for(n);
        total = std::max(total, voff[v] + n);
    }
    // volumes must not overlap
    {
        std::vector<std::pair<int64_t, int64_t>> iv(nvol);
        for (int v = 0; v < nvol; ++v)
            iv[v] = {voff[v], voff[v] + vdims[3 * v] * vdims[3 * v + 1] * vdims[3 * v + 2]};
        std::sort(iv.begin(), iv.end());
        for (int v = 1; v < nvol; ++v) MH_REQUIRE(ctx, iv[v].first >= iv[v - 1].second, "volumes overlap");
    }
    MH_REQUIRE(ctx, nblocks < (int64_t(1) << 31), "too many blocks");
    MH_REQUIRE(ctx, work_bytes >= ragged_layout(total, nvol, nblocks, kCritLogMax / 16, nullptr, nullptr),
               "workspace too small");
    MH_REQUIRE(ctx, (reinterpret_cast<uintptr_t>(work) & 255) == 0, "workspace must be 256-byte aligned");
    RaggedWorkspace W;
    ragged_layout(total, nvol, nblocks, maxcheck, static_cast<char *>(work), &W);
    hipStream_t s = (hipStream_t)stream;
    MH_CHECK(ctx, hipMemcpyAsync(W.vols, vols.data(), sizeof(Vol) * nvol, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(zero_ints, dim3((3 * nvol + 255) / 256), dim3(256), 0, s, W.stop, 3 * nvol);
    hipLaunchKernelGGL(ragged_init_kernel, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, W.vols, nvol,
                       W.q4, eigs, W.hp, W.pt, W.tmp, u, cs, ct, prm->rcc);
    MH_CHECK_LAUNCH(ctx, "ragged init");
    const PParams PP = p_params(prm);
    const UParams UP{prm->rcc, float(1.0 / double(prm->rcc)), prm->rs[0], prm->rs[1], prm->rs[2], prm->error_bound, 0.f};
    const BatchCtrl C{W.stop, W.ncrit, W.count, W.crit_log, maxcheck};
    for (int it = 0; it < prm->num_iter; ++it) {
        hipLaunchKernelGGL(ragged_p_kernel, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, W.vols, nvol,
                           W.hp, W.tmp, W.q4, alpha, PP, W.stop);
        if ((it % prm->update) == 0) {
            hipLaunchKernelGGL(ragged_u_kernel<true>, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, W.vols,
                               nvol, W.hp, W.tmp, cs, ct, W.pt, u, UP, C, W.partials);
            hipLaunchKernelGGL(ragged_crit_kernel, dim3(nvol), dim3(mh::kBlock), 0, s, W.vols, W.partials,
                               prm->rcc, prm->error_bound, C);
        } else
            hipLaunchKernelGGL(ragged_u_kernel<false>, dim3(unsigned(nblocks)), dim3(mh::kBlock), 0, s, W.vols,
                               nvol, W.hp, W.tmp, cs, ct, W.pt, u, UP, C, W.partials);
        MH_CHECK_LAUNCH(ctx, "ragged cmf kernels");
    }
    std::vector<int> st(2 * nvol);
    MH_CHECK(ctx, hipMemcpyAsync(st.data(), W.stop, sizeof(int) * 2 * nvol, hipMemcpyDeviceToHost, s));
    std::vector<float> log(int64_t(nvol) * maxcheck);
    MH_CHECK(ctx, hipMemcpyAsync(log.data(), W.crit_log, sizeof(float) * log.size(), hipMemcpyDeviceToHost, s));
    MH_CHECK(ctx, hipStreamSynchronize(s));
    int any_nan = 0;
    for (int v = 0; v < nvol; ++v) {
        const int sv = st[v], nc = std::min(st[nvol + v], maxcheck);
        any_nan |= sv == 1;
        if (res) {
            res[v].stopped = sv;
            res[v].ncrit = nc;
            res[v].iterations = sv ? (st[nvol + v] - 1) * prm->update + 1 : prm->num_iter;
            res[v].last_crit = nc > 0 ? log[int64_t(v) * maxcheck + nc - 1] : 0.f;
        }
    }
    return any_nan ? MH_STOPPED_NAN : MH_OK;
}

MH_EXPORT int mh_cmf_batched_set_schedule(mh_ctx *ctx, int group, int visit)
{
    MH_REQUIRE(ctx, group >= 0 && group <= kSI && visit >= 0, "group in 0..8, visit >= 0");
    ctx->batch_group = group;
    ctx->batch_visit = visit;
    return MH_OK;
}

MH_EXPORT int mh_cmf_set_sweep(mh_ctx *ctx, int mode, int tile, int zchunk)
{
    MH_REQUIRE(ctx, mode >= 0 && mode <= 2, "mode 0 (automatic), 1 (off) or 2 (on)");
    MH_REQUIRE(ctx, tile >= 0 && tile <= kNumSweepShapes && zchunk >= 0, "bad tile or zchunk");
    ctx->sweep_mode = mode;
    ctx->sweep_shape = tile;
    ctx->sweep_zchunk = zchunk;
    return MH_OK;
}

MH_EXPORT int mh_cmf_batched_set_sweep(mh_ctx *ctx, int mode)
{
    MH_REQUIRE(ctx, mode >= 0 && mode <= 2, "mode 0 (automatic), 1 (off) or 2 (on)");
    ctx->batch_sweep = mode;
    return MH_OK;
}

MH_EXPORT int mh_cmf_sweep_choice(mh_ctx *ctx, const int64_t dims[3], int *tile, int *zchunk)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims) && tile && zchunk, "bad arguments");
    auto f = ctx->sweep_tuned.find({int(dims[0]), int(dims[1]), int(dims[2])});
    if (f == ctx->sweep_tuned.end()) return 0;
    *tile = f->second.first + 1;  // 0: the per-iteration kernels
    *zchunk = f->second.second;
    return 1;
}

MH_EXPORT int mh_cmf_batched_sweep_choice(mh_ctx *ctx, const int64_t dims[3], int *tile, int *zchunk)
{
    MH_REQUIRE(ctx, mh::dims_ok(dims) && tile && zchunk, "bad arguments");
    auto f = ctx->batch_tuned.find({int(dims[0]), int(dims[1]), int(dims[2])});
    if (f == ctx->batch_tuned.end()) return 0;
    *tile = f->second.first + 1;
    *zchunk = f->second.second;
    return 1;
}

MH_EXPORT int mh_cmf_set_graphs(mh_ctx *ctx, int on)
{
    MH_REQUIRE(ctx, on == 0 || on == 1, "0 (off) or 1 (on)");
    ctx->graph_mode = on;
    return MH_OK;
}
