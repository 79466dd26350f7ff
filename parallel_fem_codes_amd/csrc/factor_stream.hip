// Streamed ILU(0) numeric factorisation: one workgroup walks the lower sweep's level schedule, one
// step (at most kStreamThreads rows of one level) per LDS barrier, thread t factoring row t of the
// step. A row needs the finished U part and inverse pivot of each row it eliminates with; those live
// in an LDS ring (kFactorRing values per slot, the same slots as the lower sweep's ring), so a row
// reads its dependencies from LDS only. The row's own data -- values padded to kFactorRow entries,
// the column-match codes of its (entry, L entry) pairs, its length and L count, the ring slots of
// its L entries -- is in schedule order and fetched into registers D steps ahead; the results are
// written in schedule order and scattered to A's layout afterwards by a full-device kernel. Values
// and results are stored entry-major (entry e of position q at e*n + q), so each load or store of a
// step's entry is one contiguous run across the threads: one workgroup issues all of the kernel's
// memory traffic, and a strided layout costs it a cache line per lane.
//
// Arithmetic (ILUp_precond_setup.c:84-166, p = 0; contraction off): entries in column order; entry e
// starts from a_e and subtracts l_t * u_{j_t, col_e} for the L entries t < e in order; an L entry
// then becomes the multiplier l_e = v_e * Dinv[j_e]. For each entry this is the serial IKJ
// elimination's sequence of operations, so the factors are bit-identical.
#include "common.h"
#include "device_util.h"

namespace fg {

constexpr int kFactorSink = kFactorRow + 1;   // sink planes: row entries and pivot, kStreamThreads doubles each

struct FactorArgs {
    int n, nsteps, W;
    const int *stepq;
    const unsigned short *slot;   // lower sweep plan: ring slots of the L entries (4 per position)
    const uint4 *codes;
    const double *aval;
    double *out, *dinv, *sink;
    int *err;
};

struct FactorStage {
    double a[kFactorRow];
    uint4 codes;
    unsigned short s[4];
    int len, nL, q;
};

__device__ __forceinline__ int factor_pos(const int *sq, int s)
{
    const int q = sq[s] + static_cast<int>(threadIdx.x);
    return q < sq[s + 1] ? q : -1;
}

__device__ __forceinline__ void factor_fetch(FactorStage &g, const FactorArgs &A, const int *sq, int s)
{
    const int q = factor_pos(sq, s);
    const int qq = q < 0 ? 0 : q;
#pragma unroll
    for (int e = 0; e < kFactorRow; ++e) g.a[e] = A.aval[static_cast<size_t>(e) * A.n + qq];
    g.codes = A.codes[qq];
    g.len = static_cast<int>(g.codes.x & 15u);        // bits 0-7 (the pairs of entry 0, which has none):
    g.nL = static_cast<int>((g.codes.x >> 4) & 15u);  // the row's length and L count
    const float2 sl = *reinterpret_cast<const float2 *>(A.slot + static_cast<size_t>(qq) * 4);
    __builtin_memcpy(g.s, &sl, 8);
    g.q = q;
}

__device__ __forceinline__ unsigned code_of(const uint4 &c, int e, int t)
{
    const int bit = 4 * (4 * e + t);
    const unsigned w = bit < 32 ? c.x : bit < 64 ? c.y : bit < 96 ? c.z : c.w;
    return (w >> (bit % 32)) & 15u;
}

template <int D>
__global__ __launch_bounds__(kStreamThreads) void k_factor_stream(FactorArgs A)
{
#pragma clang fp contract(off)
    static_assert(kStreamPad % D == 0, "step padding must cover the lookahead");
    extern __shared__ unsigned char smem_raw[];
    double *ring = reinterpret_cast<double *>(smem_raw);          // (W + 2) slots of kFactorRing
    int *sq = reinterpret_cast<int *>(ring + kFactorRing * (A.W + 2));
    const int nsq = A.nsteps + 2 * kStreamPad + 1;
    for (int i = threadIdx.x; i < nsq; i += blockDim.x) sq[i] = A.stepq[i];
    if (threadIdx.x < kFactorRing) ring[kFactorRing * A.W + threadIdx.x] = 0.0;   // slot W: zeros
    __syncthreads();
    const int wmask = A.W - 1;
    double *sink = A.sink + threadIdx.x;
    FactorStage st[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
#pragma unroll
        for (int e = 0; e < kFactorRow; ++e) st[u].a[e] = 0.0;
        st[u].codes = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
#pragma unroll
        for (int t = 0; t < 4; ++t) st[u].s[t] = static_cast<unsigned short>(A.W);
        st[u].len = 1;
        st[u].nL = 0;
        st[u].q = -1;
    }
    for (int s0 = 0; s0 < A.nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            __builtin_amdgcn_sched_barrier(0);
            FactorStage &g = st[u];
            // all ring reads first (unconditional, absent pairs read the zero slot), so they are
            // issued back to back; then the elimination chain with selects
            double dj[4], l[4], out[kFactorRow], r[kFactorRow][4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                dj[t] = ring[kFactorRing * g.s[t] + kFactorU];
                l[t] = 0.0;
            }
#pragma unroll
            for (int e = 1; e < kFactorRow; ++e)
#pragma unroll
                for (int t = 0; t < 4 && t < e; ++t) {
                    const unsigned code = code_of(g.codes, e, t);
                    const int at = code < static_cast<unsigned>(kFactorU) ? kFactorRing * g.s[t] + static_cast<int>(code)
                                                                          : kFactorRing * A.W;
                    r[e][t] = ring[at];
                }
            double piv = 0.0;
#pragma unroll
            for (int e = 0; e < kFactorRow; ++e) {
                double acc = g.a[e];
#pragma unroll
                for (int t = 0; t < 4 && t < e; ++t) {
                    const bool has = code_of(g.codes, e, t) < static_cast<unsigned>(kFactorU);
                    const double upd = acc - l[t] * r[e][t];
                    acc = has ? upd : acc;
                }
                if (e < 4 && e < g.nL) {
                    l[e] = acc * dj[e];
                    out[e] = l[e];
                } else {
                    out[e] = acc;
                }
                if (e == g.nL) piv = acc;
            }
            const bool on = g.q >= 0;
            double dinv = 1.0 / piv;
            if (piv == 0.0) {
                dinv = 0.0;
                if (on) __hip_atomic_store(A.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const int self = on ? (g.q & wmask) : A.W + 1;
#pragma unroll
            for (int e = 1; e < kFactorRow; ++e) {
                const int uo = e - g.nL - 1;
                if (uo >= 0 && e < g.len) ring[kFactorRing * self + uo] = out[e];
            }
            ring[kFactorRing * self + kFactorU] = dinv;
            double *o = on ? A.out + g.q : sink;
            const size_t plane = on ? static_cast<size_t>(A.n) : static_cast<size_t>(kStreamThreads);
#pragma unroll
            for (int e = 0; e < kFactorRow; ++e) o[e * plane] = out[e];
            *(on ? A.dinv + g.q : sink + kFactorRow * kStreamThreads) = dinv;
            factor_fetch(g, A, sq, s0 + u + D);
            lds_barrier();
        }
    }
}

// aval[e*n + q] = A's value e of the row at schedule position q (0 past the row's end).
__global__ void k_factor_gather(int n, const int4 *__restrict__ meta, const double *__restrict__ AA,
                                double *__restrict__ aval)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    // consecutive threads take consecutive entries of a row: the reads of A's row are contiguous
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];
    aval[static_cast<size_t>(e) * n + q] = e < m.y ? AA[m.x + e] : 0.0;
}

// The factored rows into A's layout, and the inverse pivots by row.
__global__ void k_factor_scatter(int n, const int4 *__restrict__ meta, const double *__restrict__ out,
                                 const double *__restrict__ dq, double *__restrict__ LU, double *__restrict__ Dinv)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];
    if (e < m.y) LU[m.x + e] = out[static_cast<size_t>(e) * n + q];
    if (e == 0) Dinv[m.w] = dq[q];
}

int launch_stream_factor(femgpu_ctx *ctx, femgpu_csr *A)
{
    StreamFactor &F = A->sfac;
    const StreamSweep &P = A->slo;
    const int n = A->n;
    const size_t tot = static_cast<size_t>(n) * kFactorRow;
    const int grid = static_cast<int>((tot + 255) / 256);
    FG_TRY(F.sink.alloc(static_cast<size_t>(kStreamThreads) * kFactorSink));
    k_factor_gather<<<grid, 256, 0, ctx->stream>>>(n, F.meta.p, A->AA.p, F.aval.p);
    FactorArgs a;
    a.n = n;
    a.nsteps = P.nsteps;
    a.W = P.W;
    a.stepq = P.stepq.p;
    a.slot = P.slot.p;
    a.codes = F.codes.p;
    a.aval = F.aval.p;
    a.out = F.out.p;
    a.dinv = F.dinv.p;
    a.sink = F.sink.p;
    a.err = ctx->err.p;
    k_factor_stream<4><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
    k_factor_scatter<<<grid, 256, 0, ctx->stream>>>(n, F.meta.p, F.out.p, F.dinv.p, A->LU.p, A->Dinv.p);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

}  // namespace fg
