// Wavefront/workgroup reductions for 64-wide CDNA wavefronts.
// Block sums are formed in a fixed order (xor-butterfly inside the
// wavefront, then wavefront partials in index order), so results are
// deterministic from run to run.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Result valid in thread 0.  sh: __shared__ double[BS/64].
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < BS / 64; q++) r += sh[q];
    }
    __syncthreads();
    return r;
}

template <int BS>
__device__ __forceinline__ double block_max(double v, double *sh)
{
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 1; q < BS / 64; q++) r = fmax(r, sh[q]);
    }
    __syncthreads();
    return r;
}
