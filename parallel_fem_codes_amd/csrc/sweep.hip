// Level-scheduled ILU(0) triangular sweeps, second design: LDS-staged row chunks with mailbox hand-off.
//
// The sweep order (ascending rows for L, descending for U) is cut into chunks of R consecutive rows;
// one workgroup per chunk, chunks claimed in sweep order from a ticket. Inside a chunk, rows are
// scheduled by dependency level (groups of equal level, a barrier between groups). At plan time
// every dependency of a row is encoded either as the LDS slot of a row of the same chunk, or as a
// remote row. A workgroup first stages everything its chunk needs (right-hand side, packed factor
// entries, inverse pivots) into LDS -- without waiting on anybody, so the staging of all resident
// chunks overlaps the sweep front -- and only then walks its groups.
//
// Remote values travel through per-row mailboxes of 64-bit words {epoch:32 | payload:32} (one word
// for FP32 rows, two for FP64 rows), written and polled with relaxed agent-scope atomics. Because
// the data word carries its own validity (the epoch of the current launch), no fences, flags or
// resets are needed, and every dependency is waited on individually. The epoch advances once per
// launch (by the last chunk to finish, which also rewinds the ticket), so launches need no reset
// kernel or memset in between. Waits are bounded: on expiry the sticky error flag is raised.
//
// Arithmetic per row follows ILUp_precond.c:26-44 exactly (contraction off): the result is the
// serial sweep's, bit for bit.
#include <climits>

#include "common.h"
#include "device_util.h"

namespace fg {

struct PackedArgs {
    int n, R, nchunks, emax;
    const int *chunk_ptr, *row_of, *ent_ptr, *ent_code, *grp_ptr, *grp_start, *grp_flags;   // grp_start: (start, end) pairs
    const unsigned char *exported;
    unsigned *ctl;                 // [0] ticket [1] done [2] FP32 epoch [3] FP64 launch parity counter
    unsigned long long *mbox;      // per row: 1 (FP32) or 2 (FP64) words
    int *err;
    unsigned spin_limit;
};

template <typename T>
struct MailboxIO;

template <>
struct MailboxIO<float> {
    static __device__ __forceinline__ void put(unsigned long long *mb, int row, unsigned epoch, float v)
    {
        const unsigned long long w = (static_cast<unsigned long long>(epoch) << 32) | __float_as_uint(v);
        __hip_atomic_store(mb + row, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// FP64 rows: one word holding the double itself; "not yet written" is a sentinel NaN pattern that
// no arithmetic produces. Two mailbox arrays alternate between launches (epoch parity): a launch
// reads and writes the active one and returns its own exported rows of the idle one to the
// sentinel, so the next launch finds them reset without a separate pass.
constexpr unsigned long long kSentinel = 0xfff8f00df00df00dull;

template <>
struct MailboxIO<double> {
    static __device__ __forceinline__ unsigned long long *active(const PackedArgs &S, unsigned epoch)
    {
        return S.mbox + (epoch & 1u) * static_cast<size_t>(S.n);
    }
    static __device__ __forceinline__ void put(unsigned long long *mb, int row, unsigned epoch, double v)
    {
        (void)epoch;
        __hip_atomic_store(mb + row, static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Poll the mailboxes of up to N remote dependencies together (one memory round trip per try).
template <typename T, int N>
struct RemoteGather;

template <int N>
struct RemoteGather<double, N> {
    static __device__ __forceinline__ void run(const PackedArgs &S, unsigned epoch, const int *rows, int cnt, double *out)
    {
        const unsigned long long *mb = MailboxIO<double>::active(S, epoch);
        unsigned long long w[N];
        unsigned spins = 0;
        unsigned pending = (cnt >= 32) ? 0xffffffffu : ((1u << cnt) - 1u);
        while (pending) {
#pragma unroll
            for (int t = 0; t < N; ++t)
                if (pending & (1u << t)) w[t] = __hip_atomic_load(mb + rows[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int t = 0; t < N; ++t)
                if ((pending & (1u << t)) && w[t] != kSentinel) {
                    out[t] = __longlong_as_double(static_cast<long long>(w[t]));
                    pending &= ~(1u << t);
                }
            if (!pending) break;
            if (++spins > S.spin_limit ||
                ((spins & 255u) == 0u && __hip_atomic_load(S.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(S.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                for (int t = 0; t < N; ++t) if (pending & (1u << t)) out[t] = 0.0;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
};

template <int N>
struct RemoteGather<float, N> {
    static __device__ __forceinline__ void run(const PackedArgs &S, unsigned epoch, const int *rows, int cnt, float *out)
    {
        unsigned long long w[N];
        unsigned spins = 0;
        unsigned pending = (cnt >= 32) ? 0xffffffffu : ((1u << cnt) - 1u);
        while (pending) {
#pragma unroll
            for (int t = 0; t < N; ++t)
                if (pending & (1u << t)) w[t] = __hip_atomic_load(S.mbox + rows[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int t = 0; t < N; ++t)
                if ((pending & (1u << t)) && static_cast<unsigned>(w[t] >> 32) == epoch) {
                    out[t] = __uint_as_float(static_cast<unsigned>(w[t]));
                    pending &= ~(1u << t);
                }
            if (!pending) break;
            if (++spins > S.spin_limit ||
                ((spins & 255u) == 0u && __hip_atomic_load(S.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(S.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                for (int t = 0; t < N; ++t) if (pending & (1u << t)) out[t] = 0.0f;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
};

constexpr int kPackedThreads = 512;

// UPPER: x = (b - U x) * Dinv; else x = b - L x. b may alias x (each row reads its own b during
// staging, before any row of the launch is written, and other chunks never read x: they read the
// mailboxes).
template <typename T, bool UPPER>
__global__ __launch_bounds__(kPackedThreads) void k_sweep_packed(PackedArgs S, const T *__restrict__ val,
                                                                 const T *__restrict__ dinv, const T *b, T *x,
                                                                 const int *__restrict__ skip)
{
#pragma clang fp contract(off)
    if (skip && *skip) return;   // uniform: written by an earlier launch
    extern __shared__ unsigned char smem_raw[];
    __shared__ int s_chunk;
    __shared__ unsigned s_epoch;
    const int R = S.R;
    T *xs = reinterpret_cast<T *>(smem_raw);   // results of the chunk's rows
    T *bs = xs + R;                            // right-hand side
    T *ds = bs + R;                            // inverse pivots (upper)
    T *vs = ds + R;                            // packed factor values
    int *cs = reinterpret_cast<int *>(vs + S.emax);   // packed dependency codes
    int *os = cs + S.emax;                             // R+1 entry offsets
    int *rs = os + R + 1;                              // row index, bit 31: exported
    int *gs = rs + R;                                  // group ranges of the chunk (2 per group)
    int *gf = gs + 2 * R;                              // group flags (bit 0: reads earlier chunks)
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(&S.ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_epoch = __hip_atomic_load(&S.ctl[sizeof(T) == 8 ? 3 : 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t >= static_cast<unsigned>(S.nchunks)) {
            __hip_atomic_store(S.err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_chunk = -1;
        } else {
            s_chunk = static_cast<int>(t);
        }
    }
    __syncthreads();
    const int c = s_chunk;
    if (c < 0) return;
    const unsigned epoch = s_epoch;
    unsigned long long *mbw = sizeof(T) == 8 ? S.mbox + (epoch & 1u) * static_cast<size_t>(S.n) : S.mbox;
    const int q0 = S.chunk_ptr[c], q1 = S.chunk_ptr[c + 1], rows = q1 - q0;
    const int e0 = S.ent_ptr[q0], e1 = S.ent_ptr[q1];
    // ---- staging (no waiting on other chunks)
    for (int k = threadIdx.x; k < rows; k += blockDim.x) {
        const int row = S.row_of[q0 + k];
        bs[k] = b[row];
        if (UPPER) ds[k] = dinv[q0 + k];
        os[k] = S.ent_ptr[q0 + k] - e0;
        rs[k] = row | (S.exported[q0 + k] ? int(0x80000000u) : 0);
    }
    if (threadIdx.x == 0) os[rows] = e1 - e0;
    const int g0 = S.grp_ptr[c], ng = S.grp_ptr[c + 1] - g0;
    for (int g = threadIdx.x; g < 2 * ng; g += blockDim.x) gs[g] = S.grp_start[2 * g0 + g];
    for (int g = threadIdx.x; g < ng; g += blockDim.x) gf[g] = S.grp_flags[g0 + g];
    for (int e = threadIdx.x; e < e1 - e0; e += blockDim.x) {
        cs[e] = S.ent_code[e0 + e];
        vs[e] = val[e0 + e];
    }
    __syncthreads();
    // ---- level groups: LDS reads, global stores left in flight
    for (int g = 0; g < ng; ++g) {
        const int k0 = gs[2 * g], k1 = gs[2 * g + 1];
        for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
            T acc = bs[k];
            const int pb = os[k], pe = os[k + 1];
            if (gf[g]) {
                // rows that read earlier chunks: fetch every remote value in one polling loop,
                // then accumulate in the reference's order
                constexpr int N = 8;
                for (int p = pb; p < pe; p += N) {
                    const int cnt = min(N, pe - p);
                    int rrows[N];
                    int nr = 0;
                    T rv[N];
#pragma unroll
                    for (int t = 0; t < N; ++t)
                        if (t < cnt && cs[p + t] < 0) rrows[nr++] = -cs[p + t] - 1;
                    if (nr) RemoteGather<T, N>::run(S, epoch, rrows, nr, rv);
                    int r = 0;
#pragma unroll
                    for (int t = 0; t < N; ++t) {
                        if (t >= cnt) break;
                        const int code = cs[p + t];
                        const T xv = code >= 0 ? xs[code] : rv[r++];
                        acc -= xv * vs[p + t];
                    }
                }
            } else {
                for (int p = pb; p < pe; ++p) acc -= xs[cs[p]] * vs[p];
            }
            if (UPPER) acc *= ds[k];
            xs[k] = acc;
            const int rw = rs[k];
            const int row = rw & 0x7fffffff;
            x[row] = acc;
            if (rw < 0) MailboxIO<T>::put(mbw, row, epoch, acc);
        }
        lds_barrier();
    }
    if (sizeof(T) == 8) {
        unsigned long long *idle = S.mbox + ((epoch + 1u) & 1u) * static_cast<size_t>(S.n);
        for (int k = threadIdx.x; k < rows; k += blockDim.x)
            if (rs[k] < 0) idle[rs[k] & 0x7fffffff] = kSentinel;
    }
    // ---- last chunk out rewinds the ticket and advances the epoch for the next launch
    if (threadIdx.x == 0) {
        const unsigned done = __hip_atomic_fetch_add(&S.ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (done == static_cast<unsigned>(S.nchunks) - 1u) {
            __hip_atomic_store(&S.ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&S.ctl[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // FP32 tags skip 0 (the mailboxes' initial content); FP64 only uses the parity
            unsigned ne = epoch + 1u;
            if (sizeof(T) == 4 && ne == 0u) ne = 1u;
            __hip_atomic_store(&S.ctl[sizeof(T) == 8 ? 3 : 2], ne, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}


// ------------------------------------------------------------------------------------------------
// Streamed sweep: one workgroup walks the whole level schedule, one step (at most kStreamThreads
// rows of one level) per LDS barrier, thread t taking row t of the step. Results of recent rows live
// in an LDS ring, so a row reads its dependencies from LDS only and the per-level cost is one LDS
// round trip plus one barrier -- no cross-workgroup hand-off. The vectors are in schedule order
// (permuted by full-device gathers before and after the pair of sweeps), so every load and store of
// a step is contiguous; a single CU's memory pipeline is what bounds a step, and scattered row
// accesses would multiply the cache lines it has to move. The row data of a step (padded factor
// entries, ring slots, inverse pivot, right-hand side) does not depend on any result and is fetched
// into registers D steps ahead; those loads and the result stores stay in flight across the
// LDS-only barriers. Every thread issues the same loads on every step (idle threads read position 0),
// so the compiler's wait counts see one path and wait only for the step being computed; idle lanes
// of a wavefront with rows store to a sink, and a wavefront without rows skips its one store.

struct StreamArgs {
    int nsteps;          // kStreamPad empty steps, the schedule, empty steps up to a multiple of
                         // kStreamPad; stepq has nsteps + 2*kStreamPad + 1 entries
    int W, L;            // ring entries, long-lived slots after them
    const int *stepq;
    const unsigned short *slot;
    const unsigned short *keep;   // L > 0: second write slot of each position (long-lived or W + 1)
    void *sink;          // kStreamThreads results of idle threads
};

template <typename T, int K>
struct StreamStage {
    T v[K];
    unsigned short s[K];
    T b, d;
    int q;               // schedule position of this thread's row in the staged step, -1: idle
    unsigned short ks;   // LONG: second write slot
};

template <int BYTES>
__device__ __forceinline__ void load_bytes(const void *__restrict__ src, void *dst)
{
    static_assert(BYTES % 8 == 0, "8-byte granules");
    if constexpr (BYTES % 16 == 0) {
        const float4 *p = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int i = 0; i < BYTES / 16; ++i) {
            const float4 w = p[i];
            __builtin_memcpy(static_cast<char *>(dst) + 16 * i, &w, 16);
        }
    } else {
        const float2 *p = reinterpret_cast<const float2 *>(src);
#pragma unroll
        for (int i = 0; i < BYTES / 8; ++i) {
            const float2 w = p[i];
            __builtin_memcpy(static_cast<char *>(dst) + 8 * i, &w, 8);
        }
    }
}

// Schedule position of this thread's row in step s, or -1.
__device__ __forceinline__ int stream_pos(const int *sq, int s)
{
    const int q = sq[s] + static_cast<int>(threadIdx.x);
    return q < sq[s + 1] ? q : -1;
}

// Stage step s into g: entries, slots, pivot and right-hand side by schedule position.
template <typename T, bool UPPER, int K, bool LONG>
__device__ __forceinline__ void stream_fetch(StreamStage<T, K> &g, const StreamArgs &a, const int *sq, int s,
                                             const T *__restrict__ val, const T *__restrict__ dinv, const T *b)
{
    const int q = stream_pos(sq, s);
    const int qq = q < 0 ? 0 : q;
    load_bytes<K * sizeof(T)>(val + static_cast<size_t>(qq) * K, g.v);
    load_bytes<K * 2>(a.slot + static_cast<size_t>(qq) * K, g.s);
    if (UPPER) g.d = dinv[qq];
    g.b = b[qq];
    if (LONG) g.ks = a.keep[qq];
    g.q = q;
}

// UPPER: x = (b - U x) * Dinv; else x = b - L x (ILUp_precond.c:26-44, same operation order,
// contraction off; padding entries subtract 0 * 0). b and x are in schedule order; b may alias x:
// a row's b is read D steps before its own result is stored, and no other row reads it.
template <typename T, bool UPPER, int K, int D, bool LONG>
__global__ __launch_bounds__(kStreamThreads) void k_sweep_stream(StreamArgs a, const T *__restrict__ val,
                                                                 const T *__restrict__ dinv, const T *b, T *x,
                                                                 const int *__restrict__ skip)
{
#pragma clang fp contract(off)
    static_assert(kStreamPad % D == 0 && D <= kStreamPad, "step padding must cover the lookahead");
    if (skip && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    T *ring = reinterpret_cast<T *>(smem_raw);   // W results, a zero, an idle slot, L long-lived slots
    int *sq = reinterpret_cast<int *>(ring + a.W + 2 + a.L);
    const int nsq = a.nsteps + 2 * kStreamPad + 1;
    for (int i = threadIdx.x; i < nsq; i += blockDim.x) sq[i] = a.stepq[i];
    if (threadIdx.x == 0) ring[a.W] = T(0);
    __syncthreads();
    const int wmask = a.W - 1;
    T *sink = static_cast<T *>(a.sink) + threadIdx.x;
    // The schedule starts with kStreamPad empty steps: the loop itself stages the first real
    // steps, so there is no separate prologue (whose different order of loads would make the wait
    // counts at the loop head conservative).
    StreamStage<T, K> st[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            st[u].v[k] = T(0);
            st[u].s[k] = static_cast<unsigned short>(a.W);
        }
        st[u].b = st[u].d = T(0);
        st[u].q = -1;
        st[u].ks = static_cast<unsigned short>(a.W + 1);
    }
    for (int s0 = 0; s0 < a.nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            // keep each step's work in its own region: hoisting a later step's slot decoding
            // would make this step wait for that step's loads
            __builtin_amdgcn_sched_barrier(0);
            StreamStage<T, K> &g = st[u];
            T acc = g.b;
#pragma unroll
            for (int k = 0; k < K; ++k) acc -= g.v[k] * ring[g.s[k]];
            if (UPPER) acc *= g.d;
            const bool on = g.q >= 0;
            ring[on ? (g.q & wmask) : a.W + 1] = acc;
            if (LONG) ring[g.ks] = acc;   // read again beyond the ring's reach (else the idle slot)
            // a wave past the step's rows (its lane 0 idle) stores nothing
            if (__builtin_amdgcn_readfirstlane(on ? 1 : 0)) *(on ? x + g.q : sink) = acc;
            stream_fetch<T, UPPER, K, LONG>(g, a, sq, s0 + u + D, val, dinv, b);
            lds_barrier();
        }
    }
}

// dst[i] = src[idx[i]]: the permutations into and out of schedule order.
template <typename T>
__global__ void k_permute(int n, const int *__restrict__ idx, const T *__restrict__ src, T *__restrict__ dst,
                          const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// Streamed-plan values (padding 0) and inverse pivots in schedule order from the factors.
template <typename T>
__global__ void k_stream_values(size_t E, const int *__restrict__ src, const double *__restrict__ LU,
                                T *__restrict__ val, int n, const int *__restrict__ row_of,
                                const double *__restrict__ Dinv, T *__restrict__ dinv)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < E; e += stride) {
        const int p = src[e];
        val[e] = p >= 0 ? static_cast<T>(LU[p]) : T(0);
    }
    if (dinv)
        for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < static_cast<size_t>(n); q += stride)
            dinv[q] = static_cast<T>(Dinv[row_of[q]]);
}

static StreamArgs stream_args(const StreamSweep &P)
{
    StreamArgs a;
    a.nsteps = P.nsteps;
    a.W = P.W;
    a.L = P.L;
    a.stepq = P.stepq.p;
    a.slot = P.slot.p;
    a.keep = P.L > 0 ? P.keep.p : nullptr;
    a.sink = P.sink.p;
    return a;
}

template <typename T>
constexpr int kD16 = sizeof(T) == 4 ? 4 : 2;

template <typename T, bool UPPER>
static int stream_launch(femgpu_ctx *ctx, const StreamSweep &P, const T *val, const T *dinv, const T *b, T *x,
                         const int *skip)
{
    const size_t lds = stream_lds_bytes(P, sizeof(T));
    const StreamArgs a = stream_args(P);
    if (P.K == 4 && P.L == 0)
        k_sweep_stream<T, UPPER, 4, 8, false><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    else if (P.K == 4)
        k_sweep_stream<T, UPPER, 4, 8, true><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    else if (P.K == 8 && P.L == 0)
        k_sweep_stream<T, UPPER, 8, 4, false><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    else if (P.K == 8)
        k_sweep_stream<T, UPPER, 8, 4, true><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    // 16-entry rows: 4 steps of lookahead fit the registers in FP32 (6.6 against 7.4 ms per apply on
    // the 513 x 513 cavity), 2 in FP64
    else if (P.L == 0)
        k_sweep_stream<T, UPPER, 16, kD16<T>, false><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    else
        k_sweep_stream<T, UPPER, 16, kD16<T>, true><<<1, kStreamThreads, lds, ctx->stream>>>(a, val, dinv, b, x, skip);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

// z = U^-1 L^-1 p with both sweeps streamed: p into lower-schedule order, lower sweep in place, into
// upper-schedule order, upper sweep in place, back to row order. p may alias z.
template <typename T>
static int stream_pair(femgpu_ctx *ctx, femgpu_csr *A, const T *b, T *x, const int *skip, const T *vlo,
                       const T *vup, const T *dup, T *t1, T *t2)
{
    const int n = A->n, grid = (n + 255) / 256;
    k_permute<T><<<grid, 256, 0, ctx->stream>>>(n, A->slo.row_of.p, b, t1, skip);
    FG_TRY((stream_launch<T, false>(ctx, A->slo, vlo, nullptr, t1, t1, skip)));
    k_permute<T><<<grid, 256, 0, ctx->stream>>>(n, A->stream_l2u.p, t1, t2, skip);
    FG_TRY((stream_launch<T, true>(ctx, A->sup, vup, dup, t2, t2, skip)));
    k_permute<T><<<grid, 256, 0, ctx->stream>>>(n, A->stream_pos_up.p, t2, x, skip);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

static bool use_stream(const femgpu_csr *A)
{
    return A->sweep_kernel == 0 && A->slo.ok && A->sup.ok;
}

// Packed factor values (and inverse pivots in sweep order) from the factors in A's layout.
template <typename T>
__global__ void k_pack_values(int E, const int *__restrict__ src, const double *__restrict__ LU, T *__restrict__ val,
                              int n, const int *__restrict__ row_of, const double *__restrict__ Dinv, T *__restrict__ dinv)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x)
        val[e] = static_cast<T>(LU[src[e]]);
    if (dinv)
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
            dinv[q] = static_cast<T>(Dinv[row_of[q]]);
}

static PackedArgs packed_args(femgpu_ctx *ctx, PackedSweep &P, bool fp32)
{
    PackedArgs a;
    a.n = P.n;
    a.R = P.R;
    a.nchunks = P.nchunks;
    a.emax = P.emax;
    a.chunk_ptr = P.chunk_ptr.p;
    a.row_of = P.row_of.p;
    a.ent_ptr = P.ent_ptr.p;
    a.ent_code = P.ent_code.p;
    a.grp_ptr = P.grp_ptr.p;
    a.grp_flags = P.grp_flags.p;
    a.grp_start = P.grp_start.p;
    a.exported = P.exported.p;
    a.ctl = P.ctl.p;
    a.mbox = fp32 ? P.mbox32.p : P.mbox.p;
    a.err = ctx->err.p;
    a.spin_limit = 1u << 22;
    return a;
}

int packed_refresh(femgpu_ctx *ctx, femgpu_csr *A)
{
    for (StreamSweep *P : {&A->slo, &A->sup}) {
        if (!P->ok) continue;
        const bool up = (P == &A->sup);
        const size_t E = static_cast<size_t>(P->n) * P->K;
        const int grid = static_cast<int>(std::max<size_t>(1, std::min<size_t>(4096, (E + 255) / 256)));
        k_stream_values<double><<<grid, 256, 0, ctx->stream>>>(E, P->src.p, A->LU.p, P->val64.p, P->n, P->row_of.p,
                                                               A->Dinv.p, up ? P->dinv64.p : nullptr);
        k_stream_values<float><<<grid, 256, 0, ctx->stream>>>(E, P->src.p, A->LU.p, P->val32.p, P->n, P->row_of.p,
                                                              A->Dinv.p, up ? P->dinv32.p : nullptr);
        FG_HIP(hipGetLastError());
    }
    for (PackedSweep *P : {&A->plo, &A->pup}) {
        const bool up = (P == &A->pup);
        const int grid = std::max(1, std::min(4096, (P->E + 255) / 256));
        k_pack_values<double><<<grid, 256, 0, ctx->stream>>>(P->E, P->ent_src.p, A->LU.p, P->val64.p, P->n, P->row_of.p,
                                                             A->Dinv.p, up ? P->dinv64.p : nullptr);
        k_pack_values<float><<<grid, 256, 0, ctx->stream>>>(P->E, P->ent_src.p, A->LU.p, P->val32.p, P->n, P->row_of.p,
                                                            A->Dinv.p, up ? P->dinv32.p : nullptr);
        FG_HIP(hipGetLastError());
    }
    return FEMGPU_OK;
}

template <typename T>
static int packed_pair(femgpu_ctx *ctx, femgpu_csr *A, const T *b, T *x, const int *skip, const T *vlo,
                       const T *vup, const T *dup)
{
    if (A->n == 0) return FEMGPU_OK;
    const size_t llo = packed_lds_bytes<T>(A->plo), lup = packed_lds_bytes<T>(A->pup);
    k_sweep_packed<T, false><<<A->plo.nchunks, kPackedThreads, llo, ctx->stream>>>(packed_args(ctx, A->plo, sizeof(T) == 4), vlo,
                                                                                     nullptr, b, x, skip);
    FG_HIP(hipGetLastError());
    k_sweep_packed<T, true><<<A->pup.nchunks, kPackedThreads, lup, ctx->stream>>>(packed_args(ctx, A->pup, sizeof(T) == 4), vup, dup,
                                                                                    x, x, skip);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int launch_sweep64(femgpu_ctx *ctx, femgpu_csr *A, const double *b, double *x, const int *skip)
{
    if (A->n == 0) return FEMGPU_OK;
    if (A->prec64_jacobi > 0) return launch_jacobi_apply64(ctx, A, A->prec64_jacobi, b, x, skip);
    if (use_stream(A))
        return stream_pair<double>(ctx, A, b, x, skip, A->slo.val64.p, A->sup.val64.p, A->sup.dinv64.p,
                                   A->stream_t64.p, A->stream_t64.p + A->n);
    return packed_pair<double>(ctx, A, b, x, skip, A->plo.val64.p, A->pup.val64.p, A->pup.dinv64.p);
}

int launch_sweep32(femgpu_ctx *ctx, femgpu_csr *A, const float *b, float *x, const int *skip)
{
    if (A->n == 0) return FEMGPU_OK;
    if (use_stream(A))
        return stream_pair<float>(ctx, A, b, x, skip, A->slo.val32.p, A->sup.val32.p, A->sup.dinv32.p,
                                  A->stream_t32.p, A->stream_t32.p + A->n);
    return packed_pair<float>(ctx, A, b, x, skip, A->plo.val32.p, A->pup.val32.p, A->pup.dinv32.p);
}

}  // namespace fg
