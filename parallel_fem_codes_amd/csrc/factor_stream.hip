// Streamed ILU(0) numeric factorisation: one workgroup walks the lower sweep's level schedule, one
// step (at most kStreamThreads rows of one level) per LDS barrier, thread t factoring row t of the
// step. A row needs the finished U part and inverse pivot of each row it eliminates with; those live
// in an LDS ring (kFactorRing values per slot, the same slots as the lower sweep's ring), so a row
// reads its dependencies from LDS only. The row's own data -- values padded to kFactorRow entries,
// the column-match codes of its (entry, L entry) pairs, its length and L count, the ring slots of
// its L entries -- is in schedule order and fetched into registers D steps ahead; the results are
// written in schedule order and scattered to A's layout afterwards by a full-device kernel. Values
// and results are stored as 16-byte pairs (entries e, e + 1 of position q at (e/2)*2n + 2q), so each
// load or store of a step's pair is one contiguous run across the threads: one workgroup issues all
// of the kernel's memory traffic, so its instruction count and a layout's cache lines per lane are
// what a step costs.
//
// Arithmetic (ILUp_precond_setup.c:84-166, p = 0; contraction off): entries in column order; entry e
// starts from a_e and subtracts l_t * u_{j_t, col_e} for the L entries t < e in order; an L entry
// then becomes the multiplier l_e = v_e * Dinv[j_e]. For each entry this is the serial IKJ
// elimination's sequence of operations, so the factors are bit-identical.
#include "common.h"
#include "device_util.h"

namespace fg {


struct FactorArgs {
    int n, nsteps, W;
    const int *stepq;
    const unsigned short *slot;   // lower sweep plan: ring slots of the L entries (4 per position)
    const uint4 *codes;
    const double *aval;   // values e, e + 1 (e even) of position q at (e / 2) * 2n + 2q: 16-byte pairs
    double *out, *dinv;   // results in the same pair layout over ostride = n + kStreamThreads positions
    int ostride;          // (idle lanes write past n); compact rows put the inverse pivot in value 7
    int *err;
};

// Two row shapes. The general one (ROW = kFactorRow, NL = 4) names each (entry e, L entry t) pair's
// U column in a 4-bit code, and the kernel forms the LDS address from the dependency's ring slot
// (read from the lower sweep's plan). The compact one (ROW = 7, NL = 3, every row at most 7 entries,
// 3 of them L and 3 U: the 7-point stencil of the transport meshes) fits one 16-byte record per row:
// x the 15 pairs' 2-bit U columns (3: none), y and z the three ring slots and the row's length and L
// count, so a row loads 72 bytes instead of 96.
// A stage holds only loaded words (and the position): anything derived from them is computed when
// the stage is used, D steps later, so no instruction waits on a load right after issuing it.
template <int ROW>
struct FactorStage {
    double a[ROW];
    uint4 codes;   // general: 4-bit codes; compact: the 16-byte record
    uint2 sl;      // general: ring slots of the L entries, 16 bits each
    int q;
};

template <int ROW>
__device__ __forceinline__ unsigned stage_lenword(const FactorStage<ROW> &g);

template <int ROW>
constexpr bool kCompact = ROW == 7;

// base[i] through a 32-bit byte offset (i * sizeof(T) < 2^32): a uniform base and a 32-bit lane
// offset, which the global loads and stores take as scalar base + vector offset
template <typename T>
__device__ __forceinline__ T &at32(T *base, unsigned i)
{
    const unsigned off = i * static_cast<unsigned>(sizeof(T));
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T &at32(const T *base, unsigned i)
{
    const unsigned off = i * static_cast<unsigned>(sizeof(T));
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

__device__ __forceinline__ int factor_pos(const int *sq, int s)
{
    const int q = sq[s] + static_cast<int>(threadIdx.x);
    return q < sq[s + 1] ? q : -1;
}

template <int ROW>
__device__ __forceinline__ void factor_fetch(FactorStage<ROW> &g, const FactorArgs &A, const int *sq, int s)
{
    const int q = factor_pos(sq, s);
    g.q = q;
    const unsigned qq = q < 0 ? 0u : static_cast<unsigned>(q);
    // uniform plane bases and one 32-bit lane offset: the loads address through scalar bases
#pragma unroll
    for (int e = 0; e < ROW; e += 2) {   // one 16-byte load per pair (a 7-entry row's last pair is padded)
        const double2 v = at32(reinterpret_cast<const double2 *>(A.aval + static_cast<size_t>(e) * A.n), qq);
        g.a[e] = v.x;
        if (e + 1 < ROW) g.a[e + 1] = v.y;
    }
    g.codes = at32(A.codes, qq);
    if constexpr (!kCompact<ROW>) g.sl = at32(reinterpret_cast<const uint2 *>(A.slot), qq);
}

// the row's length (bits 0-3) and L count (bits 4-7): compact rows keep them in z's high half,
// general rows in the code bits of entry 0 (which has no pairs)
template <int ROW>
__device__ __forceinline__ unsigned stage_lenword(const FactorStage<ROW> &g)
{
    if constexpr (kCompact<ROW>) return g.codes.z >> 16;
    else return g.codes.x;
}

__device__ __forceinline__ unsigned code_of(const uint4 &c, int e, int t)
{
    const int bit = 4 * (4 * e + t);
    const unsigned w = bit < 32 ? c.x : bit < 64 ? c.y : bit < 96 ? c.z : c.w;
    return (w >> (bit % 32)) & 15u;
}

template <int D, int ROW, int NL>
__global__ __launch_bounds__(kStreamThreads) void k_factor_stream(FactorArgs A)
{
#pragma clang fp contract(off)
    static_assert(ROW <= kFactorRow && NL <= 4, "row shape beyond the plan's codes");
    static_assert(!kCompact<ROW> || NL == 3, "the compact shape has 3 L entries");
    static_assert(kStreamPad % D == 0, "step padding must cover the lookahead");
    extern __shared__ unsigned char smem_raw[];
    double *ring = reinterpret_cast<double *>(smem_raw);          // (W + 2) slots of kFactorRing
    int *sq = reinterpret_cast<int *>(ring + kFactorRing * (A.W + 2));
    const int nsq = A.nsteps + 2 * kStreamPad + 1;
    for (int i = threadIdx.x; i < nsq; i += blockDim.x) sq[i] = A.stepq[i];
    if (threadIdx.x < kFactorRing) ring[kFactorRing * A.W + threadIdx.x] = 0.0;   // slot W: zeros
    __syncthreads();
    const int wmask = A.W - 1;
    FactorStage<ROW> st[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
#pragma unroll
        for (int e = 0; e < ROW; ++e) st[u].a[e] = 0.0;
        const unsigned w2 = static_cast<unsigned>(A.W) | (static_cast<unsigned>(A.W) << 16);
        if constexpr (kCompact<ROW>)   // no pairs, slots at the zero slot; length 1, no L
            st[u].codes = make_uint4(0xffffffffu, w2, static_cast<unsigned>(A.W) | (1u << 16), 0u);
        else   // no pairs; length 1, no L
            st[u].codes = make_uint4(0xffffff01u, 0xffffffffu, 0xffffffffu, 0xffffffffu);
        st[u].sl = make_uint2(w2, w2);
        st[u].q = -1;
    }
    for (int s0 = 0; s0 < A.nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            __builtin_amdgcn_sched_barrier(0);
            FactorStage<ROW> &g = st[u];
            const unsigned lw = stage_lenword(g);
            const int glen = static_cast<int>(lw & 15u), gnL = static_cast<int>((lw >> 4) & 15u);
            int gs[4];
            if constexpr (kCompact<ROW>) {
                gs[0] = static_cast<int>(g.codes.y & 0xffffu);
                gs[1] = static_cast<int>(g.codes.y >> 16);
                gs[2] = static_cast<int>(g.codes.z & 0xffffu);
                gs[3] = gs[2];
            } else {
                gs[0] = static_cast<int>(g.sl.x & 0xffffu);
                gs[1] = static_cast<int>(g.sl.x >> 16);
                gs[2] = static_cast<int>(g.sl.y & 0xffffu);
                gs[3] = static_cast<int>(g.sl.y >> 16);
            }
            // all ring reads first (unconditional, absent pairs read a zero), so they are issued
            // back to back; then the elimination chain with selects
            double dj[NL], l[NL], out[ROW], r[ROW][NL];
            bool has[ROW][NL];
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                dj[t] = ring[kFactorRing * gs[t] + kFactorU];
                l[t] = 0.0;
            }
#pragma unroll
            for (int e = 1; e < ROW; ++e)
#pragma unroll
                for (int t = 0; t < NL && t < e; ++t) {
                    if constexpr (kCompact<ROW>) {
                        // code 3 (no U column) reads the slot's unused fourth U word; has drops it
                        const unsigned code = (g.codes.x >> (2 * factor_pair_index(e, t))) & 3u;
                        r[e][t] = ring[kFactorRing * gs[t] + static_cast<int>(code)];
                        has[e][t] = code != 3u;
                    } else {
                        const unsigned code = code_of(g.codes, e, t);
                        has[e][t] = code < static_cast<unsigned>(kFactorU);
                        const int at = has[e][t] ? kFactorRing * gs[t] + static_cast<int>(code) : kFactorRing * A.W;
                        r[e][t] = ring[at];
                    }
                }
            double piv = 0.0;
#pragma unroll
            for (int e = 0; e < ROW; ++e) {
                double acc = g.a[e];
#pragma unroll
                for (int t = 0; t < NL && t < e; ++t) {
                    const double upd = acc - l[t] * r[e][t];
                    acc = has[e][t] ? upd : acc;
                }
                if (e < NL && e < gnL) {
                    l[e] = acc * dj[e];
                    out[e] = l[e];
                } else {
                    out[e] = acc;
                }
                if (e == gnL) piv = acc;
            }
            const bool on = g.q >= 0;
            double dinv = 1.0 / piv;
            if (piv == 0.0) {
                dinv = 0.0;
                if (on) __hip_atomic_store(A.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const int self = on ? (g.q & wmask) : A.W + 1;
#pragma unroll
            for (int e = 1; e < ROW; ++e) {
                const int uo = e - gnL - 1;
                if (uo >= 0 && e < glen) ring[kFactorRing * self + uo] = out[e];
            }
            ring[kFactorRing * self + kFactorU] = dinv;
            // every lane of a wave with rows stores (idle lanes past the n positions of each plane):
            // one uniform base per plane and a 32-bit lane offset; a wave past the step's rows (its
            // lane 0 idle) issues no stores
            const unsigned qo = on ? static_cast<unsigned>(g.q) : static_cast<unsigned>(A.n) + threadIdx.x;
            if (__builtin_amdgcn_readfirstlane(on ? 1 : 0)) {
#pragma unroll
                for (int e = 0; e < ROW; e += 2) {   // 16-byte stores; value 7 of a compact row: 1/pivot
                    const double hi = e + 1 < ROW ? out[e + 1] : dinv;
                    at32(reinterpret_cast<double2 *>(A.out + static_cast<size_t>(e) * A.ostride), qo) =
                        make_double2(out[e], hi);
                }
                if constexpr (!kCompact<ROW>) at32(A.dinv, qo) = dinv;
            }
            factor_fetch(g, A, sq, s0 + u + D);
            lds_barrier();
        }
    }
}

// aval[(e/2)*2n + 2q + e%2] = A's value e of the row at schedule position q (0 past the row's end).
__global__ void k_factor_gather(int n, const int4 *__restrict__ meta, const double *__restrict__ AA,
                                double *__restrict__ aval)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    // consecutive threads take consecutive entries of a row: the reads of A's row are contiguous
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];
    aval[static_cast<size_t>(e >> 1) * 2 * n + 2 * static_cast<size_t>(q) + (e & 1)] = e < m.y ? AA[m.x + e] : 0.0;
}

// The factored rows (pair layout over ostride positions) into A's layout, and the inverse pivots by
// row (compact: value 7 of the row).
__global__ void k_factor_scatter(int n, int ostride, int compact, const int4 *__restrict__ meta,
                                 const double *__restrict__ out, const double *__restrict__ dq, double *__restrict__ LU,
                                 double *__restrict__ Dinv)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];
    const size_t pos = static_cast<size_t>(e >> 1) * 2 * ostride + 2 * static_cast<size_t>(q) + (e & 1);
    if (e < m.y) LU[m.x + e] = out[pos];
    if (e == 7) Dinv[m.w] = compact ? out[pos] : dq[q];
}

int launch_stream_factor(femgpu_ctx *ctx, femgpu_csr *A)
{
    StreamFactor &F = A->sfac;
    const StreamSweep &P = A->slo;
    const int n = A->n;
    const size_t tot = static_cast<size_t>(n) * kFactorRow;
    const int grid = static_cast<int>((tot + 255) / 256);
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
    a.ostride = n + kStreamThreads;
    a.err = ctx->err.p;
    if (F.compact)
        k_factor_stream<4, 7, 3><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
    else
        k_factor_stream<4, kFactorRow, 4><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
    k_factor_scatter<<<grid, 256, 0, ctx->stream>>>(n, n + kStreamThreads, F.compact ? 1 : 0, F.meta.p, F.out.p,
                                                    F.dinv.p, A->LU.p, A->Dinv.p);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

}  // namespace fg
