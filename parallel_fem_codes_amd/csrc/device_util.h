// Device helpers shared by the kernels: 64-wide wave reductions, block reductions and the
// LDS-only barrier of the single-workgroup sweeps.
#pragma once

#include <hip/hip_runtime.h>

namespace fg {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in every thread of wave 0 (and returned to all lanes
// of wave 0). red must hold BS/64 entries.
template <int BS, typename T>
__device__ __forceinline__ T block_sum(T v, T *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    if (w == 0) {
        v = (l < BS / 64) ? red[l] : T(0);
        v = wave_sum(v);
    }
    return v;
}

// Barrier that orders LDS only: global stores and prefetch loads stay in flight across it, so a
// level step of the single-workgroup sweeps never waits for a memory round trip.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace fg
