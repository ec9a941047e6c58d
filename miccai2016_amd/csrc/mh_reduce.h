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

}  // namespace mh
