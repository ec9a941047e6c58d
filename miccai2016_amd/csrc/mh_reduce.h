// Work-group reductions for 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>

namespace mh {

// Reduce within one wavefront with butterfly exchanges (result in every lane).
template <class Op>
__device__ __forceinline__ float wave_reduce(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = Op::combine(v, __shfl_xor(v, off, 64));
    return v;
}

// Reduce over a work-group of up to 1024 threads; the result is valid in
// thread 0 (and returned to every thread of wave 0).
template <class Op>
__device__ __forceinline__ float block_reduce(float v)
{
    __shared__ float part[16];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    v = wave_reduce<Op>(v);
    __syncthreads();  // part[] may still be read by a previous call
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (wave == 0) {
        v = lane < nwaves ? part[lane] : Op::identity();
        v = wave_reduce<Op>(v);
    }
    return v;
}

// Sum of n values by one work-group: each thread keeps four independent
// partial sums (more loads in flight), combined in a fixed order, so the
// result depends only on n and the block size.
template <class Op>
__device__ __forceinline__ float block_sum_array(const float *__restrict__ x, int n)
{
    float a0 = Op::identity(), a1 = Op::identity(), a2 = Op::identity(), a3 = Op::identity();
    const int step = blockDim.x;
    int k = threadIdx.x;
    for (; k + 3 * step < n; k += 4 * step) {
        a0 = Op::combine(a0, x[k]);
        a1 = Op::combine(a1, x[k + step]);
        a2 = Op::combine(a2, x[k + 2 * step]);
        a3 = Op::combine(a3, x[k + 3 * step]);
    }
    for (; k < n; k += step) a0 = Op::combine(a0, x[k]);
    return block_reduce<Op>(Op::combine(Op::combine(a0, a1), Op::combine(a2, a3)));
}

}  // namespace mh
