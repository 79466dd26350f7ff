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

// XCD-aware block order. Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b mod 8), each XCD with its own L2. A launch of 8 * per blocks renumbers workgroup b as
// (b mod 8) * per + b / 8, so XCD x processes the contiguous range [x * per, (x + 1) * per) of the
// logical blocks: with a banded matrix its gathers stay mostly inside what its own L2 holds.
// Returns the logical block; the launch rounds its grid up to a multiple of 8 (xcd_grid) and
// blocks whose logical number reaches the needed count exit. A grid that is not a multiple of 8
// keeps the identity order (the renumbering is a bijection only for multiples of 8).
__device__ __forceinline__ int xcd_block()
{
    if (gridDim.x & 7u) return static_cast<int>(blockIdx.x);
    const int per = gridDim.x / 8;
    return static_cast<int>(blockIdx.x % 8) * per + static_cast<int>(blockIdx.x / 8);
}
inline int xcd_grid(int blocks) { return (blocks + 7) / 8 * 8; }

// sum of vals[src[q]] for q = p0 .. p1-1, added in list order. The first U entries are loaded
// together (positions past the end clamped to p0, their terms not added), so their loads overlap;
// the additions are those of the plain loop, in the same order.
template <int U>
__device__ __forceinline__ double gather_sum(const int *__restrict__ src, const double *__restrict__ vals, int p0, int p1)
{
    double v = 0.0;
    if (p1 <= p0) return v;
    double c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = vals[src[p0 + u < p1 ? p0 + u : p0]];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (p0 + u < p1) v += c[u];
    for (int q = p0 + U; q < p1; ++q) v += vals[src[q]];
    return v;
}

}  // namespace fg
