// Streamed ILU(0) numeric factorisation: one workgroup walks the lower sweep's level schedule, one
// step (at most kStreamThreads rows of one level) per LDS barrier, thread t factoring row t of the
// step. A row needs the finished U part and inverse pivot of each row it eliminates with; those live
// in an LDS ring (kFactorRing values per slot, the same slots as the lower sweep's ring), so a row
// reads its dependencies from LDS only. The row's own data -- values padded to kFactorRow entries,
// the column-match codes of its (entry, L entry) pairs, its length and L count, the ring slots of
// its L entries -- is in schedule order and fetched into registers D steps ahead; the results are
// written in schedule order and moved to A's layout afterwards by a full-device kernel. Values
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
    int ostride;          // (idle lanes write past n)
    int *err;
};

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

// LDS of both kernels: the ring ((W + 2) slots of kFactorRing values: slot W holds zeros, slot W + 1
// takes the writes of idle lanes) and the step table.
__device__ __forceinline__ double *factor_lds_setup(const FactorArgs &A, int *&sq)
{
    extern __shared__ unsigned char smem_raw[];
    double *ring = reinterpret_cast<double *>(smem_raw);
    sq = reinterpret_cast<int *>(ring + kFactorRing * (A.W + 2));
    const int nsq = A.nsteps + 2 * kStreamPad + 1;
    for (int i = threadIdx.x; i < nsq; i += blockDim.x) sq[i] = A.stepq[i];
    if (threadIdx.x < kFactorRing) ring[kFactorRing * A.W + threadIdx.x] = 0.0;
    __syncthreads();
    return ring;
}

// ---------------------------------------------------------------------------------------------
// General rows (at most kFactorRow entries, 4 of them L): each (entry e, L entry t) pair's U column
// in a 4-bit code, and the LDS address formed from the dependency's ring slot (read from the lower
// sweep's plan). A stage holds only loaded words (and the position): anything derived from them is
// computed when the stage is used, D steps later, so no instruction waits on a load right after
// issuing it.
struct FactorStage {
    double a[kFactorRow];
    uint4 codes;   // 4-bit codes; the code bits of entry 0 (which has no pairs) hold length and L count
    uint2 sl;      // ring slots of the L entries, 16 bits each
    int q;
};

__device__ __forceinline__ void factor_fetch(FactorStage &g, const FactorArgs &A, const int *sq, int s)
{
    const int q = factor_pos(sq, s);
    g.q = q;
    const unsigned qq = q < 0 ? 0u : static_cast<unsigned>(q);
    // uniform plane bases and one 32-bit lane offset: the loads address through scalar bases
#pragma unroll
    for (int e = 0; e < kFactorRow; e += 2) {   // one 16-byte load per pair
        const double2 v = at32(reinterpret_cast<const double2 *>(A.aval + static_cast<size_t>(e) * A.n), qq);
        g.a[e] = v.x;
        g.a[e + 1] = v.y;
    }
    g.codes = at32(A.codes, qq);
    g.sl = at32(reinterpret_cast<const uint2 *>(A.slot), qq);
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
    constexpr int ROW = kFactorRow, NL = 4;
    static_assert(kStreamPad % D == 0, "step padding must cover the lookahead");
    int *sq;
    double *ring = factor_lds_setup(A, sq);
    const int wmask = A.W - 1;
    FactorStage st[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
#pragma unroll
        for (int e = 0; e < ROW; ++e) st[u].a[e] = 0.0;
        const unsigned w2 = static_cast<unsigned>(A.W) | (static_cast<unsigned>(A.W) << 16);
        st[u].codes = make_uint4(0xffffff01u, 0xffffffffu, 0xffffffffu, 0xffffffffu);   // no pairs; length 1
        st[u].sl = make_uint2(w2, w2);
        st[u].q = -1;
    }
    for (int s0 = 0; s0 < A.nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            __builtin_amdgcn_sched_barrier(0);
            FactorStage &g = st[u];
            const unsigned lw = g.codes.x;
            const int glen = static_cast<int>(lw & 15u), gnL = static_cast<int>((lw >> 4) & 15u);
            const int gs[4] = {static_cast<int>(g.sl.x & 0xffffu), static_cast<int>(g.sl.x >> 16),
                               static_cast<int>(g.sl.y & 0xffffu), static_cast<int>(g.sl.y >> 16)};
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
                    const unsigned code = code_of(g.codes, e, t);
                    has[e][t] = code < static_cast<unsigned>(kFactorU);
                    const int at = has[e][t] ? kFactorRing * gs[t] + static_cast<int>(code) : kFactorRing * A.W;
                    r[e][t] = ring[at];
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
                for (int e = 0; e < ROW; e += 2)
                    at32(reinterpret_cast<double2 *>(A.out + static_cast<size_t>(e) * A.ostride), qo) =
                        make_double2(out[e], out[e + 1]);
                at32(A.dinv, qo) = dinv;
            }
            factor_fetch(g, A, sq, s0 + u + D);
            lds_barrier();
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Compact rows (the 7-point stencil of the transport meshes: at most 7 entries, 3 of them L and 3
// U) sit at fixed slots: L entries at 0-2, the diagonal at 3, U entries at 4-6, absent ones 0.
// Value 7 carries the row's 64-bit record: the 2-bit U column of each of the 15 (slot e, L slot t)
// pairs, t < min(3, e), at bit 2 * factor_pair_index(e, t) (3: none), and from bit 30 the schedule
// distance to each L entry's row in 11 bits (0: absent, read from the zero slot). An absent L entry
// has no U column in any pair and takes the zero slot's inverse pivot, so it stays 0 and subtracts
// nothing; an absent U entry is never named by a later row. A row loads four 16-byte pairs and
// stores two, its pivot and U part: the L multipliers are formed again by k_factor_finish once
// every pivot and U part is known, with the same operations in the same order.
constexpr int kFixedDistBits = 11;

struct FixedStage {
    double2 v[4];
    int q;
};

__device__ __forceinline__ void fixed_fetch(FixedStage &g, const FactorArgs &A, const int *sq, int s)
{
    const int q = factor_pos(sq, s);
    g.q = q;
    const unsigned qq = q < 0 ? 0u : static_cast<unsigned>(q);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        g.v[k] = at32(reinterpret_cast<const double2 *>(A.aval + static_cast<size_t>(2 * k) * A.n), qq);
}

__device__ __forceinline__ unsigned fixed_code(unsigned long long rec, int e, int t)
{
    return static_cast<unsigned>(rec >> (2 * factor_pair_index(e, t))) & 3u;
}

__device__ __forceinline__ int fixed_dist(unsigned long long rec, int t)
{
    return static_cast<int>((rec >> (30 + kFixedDistBits * t)) & ((1u << kFixedDistBits) - 1u));
}

// ALWAYS / NEVER: pair positions where every present L entry's row has / lacks the column of a
// present entry e (the plan has checked every row): the first subtract without a select, the second
// are not formed. Pairs of absent entries stay exact no-ops without the select: an absent L entry t
// has l = +0 and reads the zero slot (+0 * +0 subtracts +0), and an absent slot e starts at +0 and
// reads word 3 of its L entries' slots, which every row sets to +0 (+0 - (+-0) = +0).
template <int D, unsigned ALWAYS, unsigned NEVER>
__global__ __launch_bounds__(kStreamThreads) void k_factor_fixed(FactorArgs A)
{
#pragma clang fp contract(off)
    static_assert(kStreamPad % D == 0, "step padding must cover the lookahead");
    int *sq;
    double *ring = factor_lds_setup(A, sq);
    const int wmask = A.W - 1;
    FixedStage st[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {   // no row: zero values, no pairs, no L entries
#pragma unroll
        for (int k = 0; k < 3; ++k) st[u].v[k] = make_double2(0.0, 0.0);
        st[u].v[3] = make_double2(0.0, __longlong_as_double(0x3fffffffLL));
        st[u].q = -1;
    }
    for (int s0 = 0; s0 < A.nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            __builtin_amdgcn_sched_barrier(0);
            FixedStage &g = st[u];
            const bool on = g.q >= 0;
            // a wave past the step's rows (its lane 0 idle) computes, writes and stores nothing;
            // idle lanes of a wave with rows write to slot W + 1 and store past the n positions
            if (__builtin_amdgcn_readfirstlane(on ? 1 : 0)) {
                const unsigned long long rec = static_cast<unsigned long long>(__double_as_longlong(g.v[3].y));
                const unsigned lo = static_cast<unsigned>(rec), hi = static_cast<unsigned>(rec >> 32);
                // ring offsets of the L entries' slots, 32-bit
                unsigned base[3];
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const unsigned d = t == 0 ? (__builtin_amdgcn_alignbit(hi, lo, 30) & 0x7ffu)
                                              : __builtin_amdgcn_ubfe(hi, 9 + kFixedDistBits * (t - 1), kFixedDistBits);
                    const unsigned slot = d ? (static_cast<unsigned>(g.q) - d) & static_cast<unsigned>(wmask)
                                            : static_cast<unsigned>(A.W);
                    base[t] = kFactorRing * slot;
                }
                const double a[7] = {g.v[0].x, g.v[0].y, g.v[1].x, g.v[1].y, g.v[2].x, g.v[2].y, g.v[3].x};
                // all ring reads first, then the elimination chain with selects; code 3 reads the
                // slot's unused fourth U word, which the select drops
                double dj[3], l[3], out[7], r[7][3];
                unsigned code[7][3];
#pragma unroll
                for (int t = 0; t < 3; ++t) dj[t] = ring[base[t] + kFactorU];
#pragma unroll
                for (int e = 1; e < 7; ++e)
#pragma unroll
                    for (int t = 0; t < 3 && t < e; ++t) {
                        if ((NEVER >> factor_pair_index(e, t)) & 1u) continue;
                        code[e][t] = __builtin_amdgcn_ubfe(lo, 2 * factor_pair_index(e, t), 2);
                        r[e][t] = ring[base[t] + code[e][t]];
                    }
#pragma unroll
                for (int e = 0; e < 7; ++e) {
                    double acc = a[e];
#pragma unroll
                    for (int t = 0; t < 3 && t < e; ++t) {
                        const unsigned bit = 1u << factor_pair_index(e, t);
                        if (NEVER & bit) continue;
                        const double upd = acc - l[t] * r[e][t];
                        acc = (ALWAYS & bit) || code[e][t] != 3u ? upd : acc;
                    }
                    if (e < 3) l[e] = acc * dj[e];
                    out[e] = acc;
                }
                const double piv = out[3];
                double dinv = 1.0 / piv;
                if (piv == 0.0) {
                    dinv = 0.0;
                    if (on) __hip_atomic_store(A.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                const unsigned self = kFactorRing * static_cast<unsigned>(on ? (g.q & wmask) : A.W + 1);
                // word 3 (no U entry) holds +0: a padding slot e's pair with a present L entry t
                // reads it (code 3) and, where ALWAYS drops the select, subtracts l_t * (+0) from
                // the padding value +0, which stays +0, so no stale LDS word reaches the results
#pragma unroll
                for (int k = 0; k < 3; ++k) ring[self + k] = out[4 + k];
                ring[self + 3] = 0.0;
                ring[self + kFactorU] = dinv;
                // pivot and U part as two 16-byte pairs
                const unsigned qo = on ? static_cast<unsigned>(g.q) : static_cast<unsigned>(A.n) + threadIdx.x;
                at32(reinterpret_cast<double2 *>(A.out), qo) = make_double2(out[3], out[4]);
                at32(reinterpret_cast<double2 *>(A.out + 2 * static_cast<size_t>(A.ostride)), qo) =
                    make_double2(out[5], out[6]);
            }
            fixed_fetch(g, A, sq, s0 + u + D);
            lds_barrier();
        }
    }
}

// aval[(e/2)*2n + 2q + e%2] = A's value e of the row at schedule position q (0 past the row's end);
// compact rows: the value of fixed slot e (L entries at 0-2, diagonal at 3, U at 4-6) and the
// row's record as value 7.
__global__ void k_factor_gather(int n, int compact, const int4 *__restrict__ meta,
                                const unsigned long long *__restrict__ rec, const double *__restrict__ AA,
                                double *__restrict__ aval)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    // consecutive threads take consecutive entries of a row: the reads of A's row are contiguous
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];   // {first entry, length, L count, row}
    double v;
    if (!compact) {
        v = e < m.y ? AA[m.x + e] : 0.0;
    } else if (e == 7) {
        v = __longlong_as_double(static_cast<long long>(rec[q]));
    } else {
        const int src = e < 3 ? (e < m.z ? e : -1) : (m.z + e - 3 < m.y ? m.z + e - 3 : -1);
        v = src >= 0 ? AA[m.x + src] : 0.0;
    }
    aval[static_cast<size_t>(e >> 1) * 2 * n + 2 * static_cast<size_t>(q) + (e & 1)] = v;
}

// General rows: the factored rows (pair layout over ostride positions) into A's layout, and the
// inverse pivots by row.
__global__ void k_factor_scatter(int n, int ostride, const int4 *__restrict__ meta, const double *__restrict__ out,
                                 const double *__restrict__ dq, double *__restrict__ LU, double *__restrict__ Dinv)
{
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];
    const size_t pos = static_cast<size_t>(e >> 1) * 2 * ostride + 2 * static_cast<size_t>(q) + (e & 1);
    if (e < m.y) LU[m.x + e] = out[pos];
    if (e == 0) Dinv[m.w] = dq[q];
}

// Compact rows: the L multipliers l_e = (a_e - sum over L entries t < e with a U column of
// l_t * u_t) * Dinv[j_e], formed from the pivots and U parts k_factor_fixed wrote, in the order
// k_factor_fixed forms them (so bit for bit its values), then the row into A's layout and the
// inverse pivot by row. kFactorRow threads per schedule position, thread k writing entry k (k = 7:
// the inverse pivot), so a row's stores are contiguous; each forms the row's multipliers itself.
__global__ void k_factor_finish(int n, int ostride, const int4 *__restrict__ meta,
                                const unsigned long long *__restrict__ rec, const double *__restrict__ AA,
                                const double *__restrict__ out, double *__restrict__ LU, double *__restrict__ Dinv)
{
#pragma clang fp contract(off)
    const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (k >= static_cast<size_t>(n) * kFactorRow) return;
    const int q = static_cast<int>(k / kFactorRow), e = static_cast<int>(k % kFactorRow);
    const int4 m = meta[q];   // {first entry, length, L count, row}
    const unsigned long long rc = rec[q];
    const double2 *o0 = reinterpret_cast<const double2 *>(out);
    const double2 *o1 = reinterpret_cast<const double2 *>(out + 2 * static_cast<size_t>(ostride));
    const double2 own0 = o0[q], own1 = o1[q];
    if (e == 7) {
        Dinv[m.w] = own0.x == 0.0 ? 0.0 : 1.0 / own0.x;
        return;
    }
    if (e >= m.y) return;
    double v;
    if (e == m.z) {
        v = own0.x;
    } else if (e > m.z) {
        const int u = e - m.z - 1;
        v = u == 0 ? own0.y : u == 1 ? own1.x : own1.y;
    } else {
        // every load unconditional (an absent L entry reads the row itself), so no branch splits them
        double2 pu[3], uu[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int qd = t < m.z ? q - fixed_dist(rc, t) : q;
            pu[t] = o0[qd];
            uu[t] = o1[qd];
        }
        double l[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            double acc = AA[m.x + (f < m.z ? f : 0)];
#pragma unroll
            for (int t = 0; t < f; ++t) {
                const unsigned c = fixed_code(rc, f, t);
                const double ut = c == 0u ? pu[t].y : c == 1u ? uu[t].x : uu[t].y;
                const double upd = acc - l[t] * ut;
                acc = c != 3u ? upd : acc;
            }
            const double dj = pu[f].x == 0.0 ? 0.0 : 1.0 / pu[f].x;
            l[f] = acc * dj;
        }
        v = e == 0 ? l[0] : e == 1 ? l[1] : l[2];
    }
    LU[m.x + e] = v;
}

int launch_stream_factor(femgpu_ctx *ctx, femgpu_csr *A)
{
    StreamFactor &F = A->sfac;
    const StreamSweep &P = A->slo;
    const int n = A->n;
    const size_t tot = static_cast<size_t>(n) * kFactorRow;
    const int grid = static_cast<int>((tot + 255) / 256);
    k_factor_gather<<<grid, 256, 0, ctx->stream>>>(n, F.compact ? 1 : 0, F.meta.p, F.rec.p, A->AA.p, F.aval.p);
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
    if (F.compact) {
        if (F.structured)
            k_factor_fixed<4, kStructAlways, kStructNever><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
        else
            k_factor_fixed<4, 0u, 0u><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
        k_factor_finish<<<grid, 256, 0, ctx->stream>>>(n, n + kStreamThreads, F.meta.p, F.rec.p, A->AA.p, F.out.p,
                                                       A->LU.p, A->Dinv.p);
    } else {
        k_factor_stream<4><<<1, kStreamThreads, factor_lds_bytes(P), ctx->stream>>>(a);
        k_factor_scatter<<<grid, 256, 0, ctx->stream>>>(n, n + kStreamThreads, F.meta.p, F.out.p, F.dinv.p, A->LU.p,
                                                        A->Dinv.p);
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

}  // namespace fg
