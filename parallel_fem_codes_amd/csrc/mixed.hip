// FP32 preconditioner of the mixed-precision solver: a fine-grained parallel ILU(0) and Jacobi
// iterations for its triangular solves. Neither has a counterpart in the reference (whose ILU(0)
// is the serial IKJ elimination of ILUp_precond_setup.c:67-168 and whose sweeps are the serial
// loops of ILUp_precond.c:26-44); both approximate those operators with work that is parallel over
// all rows, which is what a latency-bound dependency chain of a few thousand levels is not.
//
// Parallel ILU(0) (Chow and Patel, "Fine-grained parallel incomplete LU factorization", SIAM J.
// Sci. Comput. 37(2), 2015): the factors are the fixed point of, for every (i, j) in the pattern,
//     l_ij = (a_ij - sum_{k<j} l_ik u_kj) / u_jj      (i > j)
//     u_ij =  a_ij - sum_{k<i} l_ik u_kj              (i <= j)
// Each sweep evaluates all right-hand sides from the previous iterate (two buffers, so the result
// does not depend on thread timing), starting from l_ij = a_ij / a_jj, u_ij = a_ij. The products of
// row i are exactly the IKJ update list of row i (the numeric factorisation plan), so the same list
// drives both the exact and the parallel factorisation.
//
// Jacobi triangular solves (Anzt, Chow and Dongarra, "Iterative sparse triangular solves for
// preconditioning", Euro-Par 2015): y = L^-1 b by y_{k+1} = b - L_s y_k from y_0 = b, and
// z = U^-1 y by z_{k+1} = Dinv (y - U_s z_k) from z_0 = Dinv y, a fixed number of iterations, so
// the preconditioner is one fixed linear operator for the Krylov method.
#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

// Initial iterate, and the inverse pivots a parallel sweep divides by.
__global__ void k_parilu_init(int n, const int *__restrict__ IA, const int *__restrict__ JA,
                              const int *__restrict__ diag, const float *__restrict__ AA, float *__restrict__ LU,
                              float *__restrict__ Dinv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = diag[i];
    for (int p = IA[i]; p < d; ++p) LU[p] = AA[p] / AA[diag[JA[p]]];
    for (int p = d; p < IA[i + 1]; ++p) LU[p] = AA[p];
    Dinv[i] = 1.0f / AA[d];
}

// One Jacobi-style sweep of the parallel ILU(0), one row per thread (rows in the plan's order, row
// order by default so neighbouring threads touch neighbouring rows; any order gives the same result). Row entries are kept in registers when the row has at most MAXROW
// entries and its list at most MAXU updates; other rows take the in-memory path.
template <int MAXROW, int MAXU>
__global__ __launch_bounds__(256) void k_parilu_sweep(int n, const int4 *__restrict__ rowinfo, const int2 *__restrict__ upd,
                                                      const int *__restrict__ JA, const float *__restrict__ AA,
                                                      const float *__restrict__ LUo, const float *__restrict__ Dio,
                                                      float *__restrict__ LUn, float *__restrict__ Din,
                                                      const int *__restrict__ row_of)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int4 ri = rowinfo[q];
    const int p0 = ri.x, len = ri.y, nL = ri.z, ub = ri.w, nu = rowinfo[q + 1].w - ub;
    const int i = row_of ? row_of[q] : q;
    float piv;
    if (len <= MAXROW && nu <= MAXU) {
        float v[MAXROW];
#pragma unroll
        for (int e = 0; e < MAXROW; ++e) v[e] = (e < len) ? AA[p0 + e] : 0.0f;
#pragma unroll
        for (int u = 0; u < MAXU; ++u) {
            if (u < nu) {
                const int2 w = upd[ub + u];
                const int t = w.y >> 16, dest = w.y & 0xffff;
                const float c = LUo[p0 + t] * LUo[w.x];
#pragma unroll
                for (int e = 0; e < MAXROW; ++e)
                    if (e == dest) v[e] -= c;
            }
        }
#pragma unroll
        for (int e = 0; e < MAXROW; ++e) {
            if (e < nL) LUn[p0 + e] = v[e] * Dio[JA[p0 + e]];
            else if (e < len) LUn[p0 + e] = v[e];
        }
        piv = LUn[p0 + nL];
    } else {
        for (int e = 0; e < len; ++e) LUn[p0 + e] = AA[p0 + e];
        for (int u = ub; u < ub + nu; ++u) {
            const int2 w = upd[u];
            const int t = w.y >> 16, dest = w.y & 0xffff;
            LUn[p0 + dest] -= LUo[p0 + t] * LUo[w.x];
        }
        for (int e = 0; e < nL; ++e) LUn[p0 + e] *= Dio[JA[p0 + e]];
        piv = LUn[p0 + nL];
    }
    Din[i] = 1.0f / piv;
}

// y_out = b - L_s y_in (y_in == nullptr: y_in = b).
template <int LPR>
__global__ __launch_bounds__(256) void k_jacobi_lower(int n, const int *__restrict__ IA, const int *__restrict__ JA,
                                                      const int *__restrict__ diag, const float *__restrict__ LU,
                                                      const float *__restrict__ b, const float *__restrict__ yin,
                                                      float *__restrict__ yout, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int gt = blockIdx.x * 256 + threadIdx.x;
    const int row = gt / LPR, lane = gt % LPR;
    if (row >= n) return;
    const float *src = yin ? yin : b;
    float acc = 0.0f;
    for (int p = IA[row] + lane; p < diag[row]; p += LPR) acc += LU[p] * src[JA[p]];
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (lane == 0) yout[row] = b[row] - acc;
}

// z_out = Dinv (y - U_s z_in) (z_in == nullptr: z_in = Dinv y).
template <int LPR>
__global__ __launch_bounds__(256) void k_jacobi_upper(int n, const int *__restrict__ IA, const int *__restrict__ JA,
                                                      const int *__restrict__ diag, const float *__restrict__ LU,
                                                      const float *__restrict__ Dinv, const float *__restrict__ y,
                                                      const float *__restrict__ zin, float *__restrict__ zout,
                                                      const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int gt = blockIdx.x * 256 + threadIdx.x;
    const int row = gt / LPR, lane = gt % LPR;
    if (row >= n) return;
    float acc = 0.0f;
    for (int p = diag[row] + 1 + lane; p < IA[row + 1]; p += LPR) {
        const int j = JA[p];
        acc += LU[p] * (zin ? zin[j] : Dinv[j] * y[j]);
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (lane == 0) zout[row] = Dinv[row] * (y[row] - acc);
}

// ---- the same iteration on the sliced-ELL layout of the FP32 matrix (slices of 64 rows, entry k of
// row 64s + l at sell_ptr[s] + 64k + l): every per-row load and store of a step is one contiguous
// run across the wave, and the update lists are stored the same way (padded per slice).
template <typename T, int MAXW>
__global__ __launch_bounds__(256) void k_parilu_init_sell(int n, const int *__restrict__ sptr,
                                                          const int *__restrict__ col, const T *__restrict__ AA,
                                                          const int *__restrict__ rinfo, const int *__restrict__ dpos,
                                                          T *__restrict__ LU, T *__restrict__ Dinv)
{
    const int i = xcd_block() * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = i >> 6, l = i & 63, p0 = sptr[s] + l, w = (sptr[s + 1] - sptr[s]) >> 6;
    const int nL = rinfo[i] & 0xff;
    // all loads of the row first (steps past the slice width re-read step 0), then the two-level
    // gathers of the column pivots
    int c[MAXW];
    T a[MAXW], piv[MAXW];
#pragma unroll
    for (int k = 0; k < MAXW; ++k) {
        const int e = p0 + ((k < w ? k : 0) << 6);
        c[k] = col[e];
        a[k] = AA[e];
    }
    const T own = AA[dpos[i]];
#pragma unroll
    for (int k = 0; k < MAXW; ++k) piv[k] = AA[dpos[c[k]]];
#pragma unroll
    for (int k = 0; k < MAXW; ++k)
        if (k < w) LU[p0 + (k << 6)] = k < nL ? a[k] / piv[k] : a[k];
    Dinv[i] = T(1) / own;
}

// One sweep. FIRST: the sweep from the initial guess (L = strict lower part of A with each column
// divided by its diagonal entry, U = upper part, pivots = the diagonal) computed from A's values
// directly -- the same divisions and products as k_parilu_init_sell followed by a sweep, without
// storing the guess. ELL: the result's strict triangles are also written to the Jacobi sliced-ELL
// arrays (vl, vu; their padding entries are zero from allocation), in place of
// k_parilu_sell_to_ell.
template <typename T, int MAXW, bool FIRST, bool ELL>
__global__ __launch_bounds__(256) void k_parilu_sweep_sell(int n, const int *__restrict__ sptr,
                                                           const int *__restrict__ col, const T *__restrict__ AA,
                                                           const int *__restrict__ rinfo,
                                                           const int *__restrict__ uptr, const int *__restrict__ usrc,
                                                           const unsigned char *__restrict__ ucode,
                                                           const int *__restrict__ pcol,
                                                           const T *__restrict__ LUo, const T *__restrict__ Dio,
                                                           T *__restrict__ LUn, T *__restrict__ Din,
                                                           const int *__restrict__ sl, const int *__restrict__ su,
                                                           T *__restrict__ vl, T *__restrict__ vu)
{
    const int i = xcd_block() * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = i >> 6, l = i & 63, p0 = sptr[s] + l, w = (sptr[s + 1] - sptr[s]) >> 6;
    const int u0 = uptr[s] + l, nu = (uptr[s + 1] - uptr[s]) >> 6;
    const int nL = rinfo[i] & 0xff;
    T v[MAXW], lo[MAXW], dcol[MAXW];
#pragma unroll
    for (int k = 0; k < MAXW; ++k) {
        v[k] = k < w ? AA[p0 + (k << 6)] : T(0);
        if (!FIRST) lo[k] = k < w ? LUo[p0 + (k << 6)] : T(0);
    }
    if (FIRST) {
        // the guess's multipliers a_ij / a_jj and pivot inverses 1 / a_jj of the lower columns
#pragma unroll
        for (int k = 0; k < MAXW; ++k) {
            const int e = p0 + ((k < w ? k : 0) << 6);
            const T piv = AA[pcol[e]];
            lo[k] = k < nL ? v[k] / piv : v[k];
            dcol[k] = T(1) / piv;
        }
    }
    // update terms in groups of four: the group's codes, sources and factor values are loaded
    // together (positions past the row's list clamped to its first, their terms skipped), then
    // applied in list order
    constexpr int G = 4;
    for (int u = 0; u < nu; u += G) {
        unsigned code[G];
        T other[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int q = u0 + ((u + g < nu ? u + g : 0) << 6);
            code[g] = u + g < nu ? ucode[q] : 0xffu;
            other[g] = FIRST ? AA[usrc[q]] : LUo[usrc[q]];   // a strict-upper entry of an earlier row
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (code[g] != 0xffu) {
                const unsigned t = code[g] >> 4, dest = code[g] & 15u;
                T lt = T(0);
#pragma unroll
                for (int k = 0; k < MAXW; ++k) lt = (k == static_cast<int>(t)) ? lo[k] : lt;
                const T c = lt * other[g];
#pragma unroll
                for (int k = 0; k < MAXW; ++k)
                    if (k == static_cast<int>(dest)) v[k] -= c;
            }
    }
    T piv = T(1);
    const int len = rinfo[i] >> 8;
    const int ql = ELL ? sl[s] + l : 0, qu = ELL ? su[s] + l : 0;
#pragma unroll
    for (int k = 0; k < MAXW; ++k)
        if (k < w) {
            const int e = p0 + (k << 6);
            const T out = k < nL ? v[k] * (FIRST ? dcol[k] : Dio[col[e]]) : v[k];
            LUn[e] = out;
            if (k == nL) piv = out;
            if (ELL) {
                if (k < nL) vl[ql + (k << 6)] = out;
                else if (k > nL && k < len) vu[qu + ((k - nL - 1) << 6)] = out;
            }
        }
    Din[i] = T(1) / piv;
}

// The FP32 and FP64 parallel factorisations: `sweeps` sweeps into the (L1, D1) / (L0, D0) ping-pong
// ending in (La, Da); the last one also fills the Jacobi arrays when vl is set.
template <typename T, int MAXW>
static void parilu_sweeps(femgpu_ctx *ctx, const femgpu_csr *A, int sweeps, const T *val, T *La, T *Lb, T *Da,
                          T *Db, T *vl, T *vu)
{
    const int n = A->n, grid = xcd_grid((n + 255) / 256);
    // sweep k writes Ln and the pair swaps: sweep 1 writes La when the count is odd, so the last
    // sweep writes La
    T *Ln = (sweeps % 2 == 1) ? La : Lb, *Lo = (sweeps % 2 == 1) ? Lb : La;
    T *Dn = (sweeps % 2 == 1) ? Da : Db, *Do = (sweeps % 2 == 1) ? Db : Da;
    for (int k = 1; k <= sweeps; ++k) {
        const bool first = k == 1, last = k == sweeps && vl;
#define FG_PARILU_SWEEP(F, E)                                                                                         \
    k_parilu_sweep_sell<T, MAXW, F, E><<<grid, 256, 0, ctx->stream>>>(                                                 \
        n, A->sell_ptr.p, A->sell_col.p, val, A->par_rinfo.p, A->par_uptr.p, A->par_usrc.p, A->par_ucode.p,            \
        A->par_pcol.p, Lo, Do, Ln, Dn, A->ell_sl.p, A->ell_su.p, vl, vu)
        if (first && last) FG_PARILU_SWEEP(true, true);
        else if (first) FG_PARILU_SWEEP(true, false);
        else if (last) FG_PARILU_SWEEP(false, true);
        else FG_PARILU_SWEEP(false, false);
#undef FG_PARILU_SWEEP
        std::swap(Lo, Ln);
        std::swap(Do, Dn);
    }
}

// The strict triangles of the SELL-layout factors into the Jacobi ELL arrays.
template <typename T>
__global__ void k_parilu_sell_to_ell(int n, const int *__restrict__ sptr, const int *__restrict__ rinfo,
                                     const T *__restrict__ LU, const int *__restrict__ sl,
                                     const int *__restrict__ su, T *__restrict__ vl, T *__restrict__ vu)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = i >> 6, l = i & 63;
    const int p0 = sptr[s] + l, nL = rinfo[i] & 0xff, len = rinfo[i] >> 8;
    const int ql = sl[s] + l, wl = (sl[s + 1] - sl[s]) >> 6, qu = su[s] + l, wu = (su[s + 1] - su[s]) >> 6;
    for (int k = 0; k < wl; ++k) vl[ql + (k << 6)] = k < nL ? LU[p0 + (k << 6)] : T(0);
    for (int k = 0; k < wu; ++k) {
        const int e = nL + 1 + k;
        vu[qu + (k << 6)] = e < len ? LU[p0 + (e << 6)] : T(0);
    }
}

// The SELL-layout factors into A's layout (LU32), for callers that read the factors back.
__global__ void k_parilu_sell_to_csr(int n, const int *__restrict__ sptr, const int *__restrict__ IA,
                                     const float *__restrict__ LU, float *__restrict__ LU32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p0 = sptr[i >> 6] + (i & 63), a0 = IA[i], len = IA[i + 1] - a0;
    for (int k = 0; k < len; ++k) LU32[a0 + k] = LU[p0 + (k << 6)];
}

int build_parilu_sell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag)
{
    A->par_ok = 0;
    const int n = A->n;
    if (!A->sell_ok || n == 0) return FEMGPU_OK;
    const std::vector<int> &sp = A->hsell_ptr;
    const int ns = static_cast<int>(sp.size()) - 1;
    int maxw = 0;
    for (int s = 0; s < ns; ++s) maxw = std::max(maxw, (sp[s + 1] - sp[s]) / 64);
    if (maxw > 16) return FEMGPU_OK;
    auto pos = [&](int row, int k) { return sp[row >> 6] + 64 * k + (row & 63); };
    std::vector<int> rinfo(n), dpos(n), slot(n, -1);
    std::vector<std::vector<std::pair<int, unsigned char>>> lists(64);
    std::vector<int> uptr(static_cast<size_t>(ns) + 1, 0), usrc;
    std::vector<unsigned char> ucode;
    for (int s = 0; s < ns; ++s) {
        size_t nu = 0;
        for (int l = 0; l < 64; ++l) {
            lists[l].clear();
            const int i = 64 * s + l;
            if (i >= n) continue;
            const int p0 = IA[i], len = IA[i + 1] - p0, nL = diag[i] - p0;
            if (nL > 15) return FEMGPU_OK;
            rinfo[i] = nL | (len << 8);
            dpos[i] = pos(i, nL);
            for (int p = p0; p < p0 + len; ++p) slot[JA[p]] = p - p0;
            for (int t = 0; t < nL; ++t) {   // the IKJ update list of row i
                const int j = JA[p0 + t];
                for (int k = diag[j] + 1; k < IA[j + 1]; ++k) {
                    const int dest = slot[JA[k]];
                    if (dest >= 0) lists[l].emplace_back(pos(j, k - IA[j]), static_cast<unsigned char>(dest | (t << 4)));
                }
            }
            for (int p = p0; p < p0 + len; ++p) slot[JA[p]] = -1;
            nu = std::max(nu, lists[l].size());
        }
        const size_t base = usrc.size();
        usrc.resize(base + 64 * nu, 0);
        ucode.resize(base + 64 * nu, 0xff);
        for (int l = 0; l < 64; ++l)
            for (size_t u = 0; u < lists[l].size(); ++u) {
                usrc[base + 64 * u + l] = lists[l][u].first;
                ucode[base + 64 * u + l] = lists[l][u].second;
            }
        if (usrc.size() >= static_cast<size_t>(INT_MAX)) return FEMGPU_OK;
        uptr[static_cast<size_t>(s) + 1] = static_cast<int>(usrc.size());
    }
    if (usrc.empty()) {   // no L entries at all: keep one padding update so the buffers exist
        usrc.push_back(0);
        ucode.push_back(0xff);
    }
    const size_t tot = A->sell_col.n;
    FG_TRY(A->par_uptr.upload(uptr.data(), uptr.size(), ctx->stream));
    FG_TRY(A->par_usrc.upload(usrc.data(), usrc.size(), ctx->stream));
    FG_TRY(A->par_ucode.upload(ucode.data(), ucode.size(), ctx->stream));
    FG_TRY(A->par_rinfo.upload(rinfo.data(), rinfo.size(), ctx->stream));
    FG_TRY(A->par_dpos.upload(dpos.data(), dpos.size(), ctx->stream));
    {   // the diagonal position of each entry's column: one gather less in the first sweep
        std::vector<int> pcol(tot, dpos[0]);
        for (int i = 0; i < n; ++i) {
            const int s = i >> 6, w = (sp[s + 1] - sp[s]) / 64, len = IA[i + 1] - IA[i];
            for (int k = 0; k < w; ++k) pcol[pos(i, k)] = k < len ? dpos[JA[IA[i] + k]] : dpos[i];
        }
        FG_TRY(A->par_pcol.upload(pcol.data(), pcol.size(), ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));   // pcol ends here
    }
    FG_TRY(A->par_a.alloc(tot));
    FG_TRY(A->par_b.alloc(tot));
    FG_HIP(hipStreamSynchronize(ctx->stream));   // the host arrays end with this call
    A->par_maxw = maxw;
    A->par_ok = 1;
    return FEMGPU_OK;
}

int launch_parilu_to_csr(femgpu_ctx *ctx, femgpu_csr *A)
{
    k_parilu_sell_to_csr<<<(A->n + 255) / 256, 256, 0, ctx->stream>>>(A->n, A->sell_ptr.p, A->IA.p, A->par_a.p,
                                                                      A->LU32.p);
    FG_HIP(hipGetLastError());
    A->lu32_in_sell = false;
    return FEMGPU_OK;
}

static int ensure_par(femgpu_csr *A)
{
    const size_t nnz = static_cast<size_t>(A->nnz) + 1, n = static_cast<size_t>(A->n) + 1;
    FG_TRY(A->LU32.alloc(nnz));
    FG_TRY(A->LU32b.alloc(nnz));
    FG_TRY(A->Dinv32.alloc(n));
    FG_TRY(A->Dinv32b.alloc(n));
    FG_TRY(A->jac_a.alloc(n));
    FG_TRY(A->jac_b.alloc(n));
    return FEMGPU_OK;
}

// The strict triangles of the FP32 factors as sliced ELL: rows in slices of 64 (one wavefront),
// each slice padded to its own longest row, entries column-major inside the slice (entry k of row
// i at sl[i / 64] + 64 k + i % 64): one thread per row, every load of a step of the row loop one
// contiguous 256-byte run, a loop count uniform across the wavefront, and padding only where a
// slice's rows differ in length. Padding entries have value 0 and column i.
constexpr int kEllMax = 16;   // widest strict part kept in sliced ELL (kernels unrolled to 4, 8 or 16)

__global__ void k_jacobi_ell_values(int n, const int *__restrict__ IA, const int *__restrict__ diag,
                                    const float *__restrict__ LU, const int *__restrict__ sl,
                                    const int *__restrict__ su, float *__restrict__ vl, float *__restrict__ vu)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = i >> 6, l = i & 63;
    const int p0 = IA[i], d = diag[i], p1 = IA[i + 1];
    const int ql = sl[s] + l, wl = (sl[s + 1] - sl[s]) >> 6, qu = su[s] + l, wu = (su[s + 1] - su[s]) >> 6;
    for (int k = 0; k < wl; ++k) vl[ql + (k << 6)] = p0 + k < d ? LU[p0 + k] : 0.0f;
    for (int k = 0; k < wu; ++k) vu[qu + (k << 6)] = d + 1 + k < p1 ? LU[d + 1 + k] : 0.0f;
}

// A block of 256 threads takes kJacRows * 256 consecutive rows, each thread kJacRows rows 256 apart
// (same lane, so a wavefront still covers whole 64-row slices): the loads of all its rows are in
// flight together, and one round of resident wavefronts covers the matrix. Every row issues all K
// column and value loads of its slice unconditionally (steps past the slice width re-read step 0,
// rows past n re-read row n-1) and only then gathers: a load behind a branch on the slice width
// would wait for the loads before it. The terms past the width are not added, so the sum is the
// one of the plain loop over the slice, in the same order.
constexpr int kJacRows = 2;

template <typename T, int K>
struct EllRows {
    int col[kJacRows][K];
    T val[kJacRows][K];
    int w[kJacRows];
};

// Column of a stored entry of row i: int columns as stored; 16-bit distances from the row (lower
// triangle i - j, upper j - i; padding entries 0, i.e. column i).
template <typename CT, bool UP>
__device__ __forceinline__ int ell_col(CT c, int i);
template <>
__device__ __forceinline__ int ell_col<int, false>(int c, int) { return c; }
template <>
__device__ __forceinline__ int ell_col<int, true>(int c, int) { return c; }
template <>
__device__ __forceinline__ int ell_col<unsigned short, false>(unsigned short c, int i) { return i - static_cast<int>(c); }
template <>
__device__ __forceinline__ int ell_col<unsigned short, true>(unsigned short c, int i) { return i + static_cast<int>(c); }

// Loads the slice entries of the thread's rows; w[r] = 0 for rows past n and for empty slices.
template <typename T, int K, typename CT, bool UP>
__device__ __forceinline__ void ell_load(int n, int blk, const int *__restrict__ sp, const CT *__restrict__ col,
                                         const T *__restrict__ val, EllRows<T, K> &e)
{
#pragma unroll
    for (int r = 0; r < kJacRows; ++r) {
        const int i = (blk * kJacRows + r) * 256 + threadIdx.x;
        const int ii = i < n ? i : n - 1;
        const int s = ii >> 6, q = sp[s] + (ii & 63), w = (sp[s + 1] - sp[s]) >> 6;
        e.w[r] = i < n ? w : 0;
        if (w > 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int p = q + ((k < w ? k : 0) << 6);
                e.col[r][k] = ell_col<CT, UP>(col[p], ii);
                e.val[r][k] = val[p];
            }
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                e.col[r][k] = ii;
                e.val[r][k] = T(0);
            }
        }
    }
}

// y_out = b - L_s y_in (y_in == nullptr: y_in = b), sliced ELL.
template <typename T, int K, typename CT>
__global__ __launch_bounds__(256) void k_jacobi_lower_ell(int n, const int *__restrict__ sl, const CT *__restrict__ cl,
                                                          const T *__restrict__ vl, const T *__restrict__ b,
                                                          const T *__restrict__ yin, T *__restrict__ yout,
                                                          const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const T *src = yin ? yin : b;
    const int blk = xcd_block();
    EllRows<T, K> e;
    ell_load<T, K, CT, false>(n, blk, sl, cl, vl, e);
    T bi[kJacRows];
#pragma unroll
    for (int r = 0; r < kJacRows; ++r) {
        const int i = (blk * kJacRows + r) * 256 + threadIdx.x;
        bi[r] = b[i < n ? i : n - 1];
    }
#pragma unroll
    for (int r = 0; r < kJacRows; ++r) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T t = e.val[r][k] * src[e.col[r][k]];
            acc = k < e.w[r] ? acc + t : acc;
        }
        const int i = (blk * kJacRows + r) * 256 + threadIdx.x;
        if (i < n) yout[i] = bi[r] - acc;
    }
}

// The first lower iteration from the FP64 residual: b = float(r / ||r||) (the products of the
// scaled FP32 copy, formed here for the row and its gathers instead of by a separate kernel), y_out =
// b - L_s b, and b stored for the next iterations.
template <int K, typename CT>
__global__ __launch_bounds__(256) void k_jacobi_lower_ell_scaled(int n, const int *__restrict__ sl,
                                                                 const CT *__restrict__ cl,
                                                                 const float *__restrict__ vl,
                                                                 const double *__restrict__ r,
                                                                 const double *__restrict__ rn,
                                                                 float *__restrict__ bout, float *__restrict__ yout)
{
    const double inv = 1.0 / *rn;
    const int blk = xcd_block();
    EllRows<float, K> e;
    ell_load<float, K, CT, false>(n, blk, sl, cl, vl, e);
    float bi[kJacRows];
#pragma unroll
    for (int q = 0; q < kJacRows; ++q) {
        const int i = (blk * kJacRows + q) * 256 + threadIdx.x;
        bi[q] = static_cast<float>(r[i < n ? i : n - 1] * inv);
    }
#pragma unroll
    for (int q = 0; q < kJacRows; ++q) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float t = e.val[q][k] * static_cast<float>(r[e.col[q][k]] * inv);
            acc = k < e.w[q] ? acc + t : acc;
        }
        const int i = (blk * kJacRows + q) * 256 + threadIdx.x;
        if (i < n) {
            yout[i] = bi[q] - acc;
            bout[i] = bi[q];
        }
    }
}

// z_out = Dinv (y - U_s z_in) (z_in == nullptr: z_in = Dinv y), sliced ELL. dpart != nullptr: the
// block's FP64 partial of <dv, z_out> (dv == nullptr: ||z_out||^2) into dpart[block], for a
// following kernel to sum over the blocks in order.
template <typename T, int K, typename CT>
__global__ __launch_bounds__(256) void k_jacobi_upper_ell(int n, const int *__restrict__ su, const CT *__restrict__ cu,
                                                          const T *__restrict__ vu, const T *__restrict__ Dinv,
                                                          const T *__restrict__ y, const T *__restrict__ zin,
                                                          T *__restrict__ zout, const int *__restrict__ skip,
                                                          const T *__restrict__ dv, double *__restrict__ dpart)
{
    __shared__ double red[4];
    if (skip && *skip) return;
    const int blk = xcd_block();
    EllRows<T, K> e;
    ell_load<T, K, CT, true>(n, blk, su, cu, vu, e);
    T di[kJacRows], yi[kJacRows];
#pragma unroll
    for (int r = 0; r < kJacRows; ++r) {
        const int i = (blk * kJacRows + r) * 256 + threadIdx.x, ii = i < n ? i : n - 1;
        di[r] = Dinv[ii];
        yi[r] = y[ii];
    }
    double dsum = 0.0;
#pragma unroll
    for (int r = 0; r < kJacRows; ++r) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = e.col[r][k];
            const T t = e.val[r][k] * (zin ? zin[j] : Dinv[j] * y[j]);
            acc = k < e.w[r] ? acc + t : acc;
        }
        const int i = (blk * kJacRows + r) * 256 + threadIdx.x;
        if (i < n) {
            const T z = di[r] * (yi[r] - acc);
            zout[i] = z;
            if (dpart) dsum += static_cast<double>(dv ? dv[i] : z) * static_cast<double>(z);
        }
    }
    if (dpart) {
        dsum = block_sum<256>(dsum, red);
        if (threadIdx.x == 0) dpart[blk] = dsum;
    }
}

int build_jacobi_ell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag)
{
    const int n = A->n;
    int kl = 0, ku = 0;
    for (int i = 0; i < n; ++i) {
        kl = std::max(kl, diag[i] - IA[i]);
        ku = std::max(ku, IA[i + 1] - diag[i] - 1);
    }
    A->ell_kl = A->ell_ku = 0;
    if (n == 0 || kl > kEllMax || ku > kEllMax) return FEMGPU_OK;
    // slice pointers: each 64-row slice padded to its own longest strict lower / upper part
    const int ns = (n + 63) / 64;
    std::vector<int> sl(static_cast<size_t>(ns) + 1, 0), su(static_cast<size_t>(ns) + 1, 0);
    for (int s = 0; s < ns; ++s) {
        int wl = 0, wu = 0;
        for (int i = 64 * s; i < std::min(n, 64 * s + 64); ++i) {
            wl = std::max(wl, diag[i] - IA[i]);
            wu = std::max(wu, IA[i + 1] - diag[i] - 1);
        }
        sl[static_cast<size_t>(s) + 1] = sl[s] + 64 * wl;
        su[static_cast<size_t>(s) + 1] = su[s] + 64 * wu;
    }
    std::vector<int> cl(std::max(sl[ns], 1)), cu(std::max(su[ns], 1));
    for (int s = 0; s < ns; ++s) {
        const int wl = (sl[s + 1] - sl[s]) / 64, wu = (su[s + 1] - su[s]) / 64;
        for (int l = 0; l < 64; ++l) {
            const int i = 64 * s + l, r = std::min(i, n - 1);
            for (int k = 0; k < wl; ++k)
                cl[static_cast<size_t>(sl[s]) + 64 * k + l] = i < n && IA[i] + k < diag[i] ? JA[IA[i] + k] : r;
            for (int k = 0; k < wu; ++k)
                cu[static_cast<size_t>(su[s]) + 64 * k + l] =
                    i < n && diag[i] + 1 + k < IA[i + 1] ? JA[diag[i] + 1 + k] : r;
        }
    }
    kl = std::max(kl, 1);
    ku = std::max(ku, 1);
    // 16-bit distances from the row when the band allows (2 bytes per entry instead of 4)
    std::vector<unsigned short> dl(cl.size()), du(cu.size());
    bool fits = true;
    for (int s = 0; s < ns && fits; ++s)
        for (int l = 0; l < 64 && fits; ++l) {
            const int r = std::min(64 * s + l, n - 1);
            for (int p = sl[s] + l; p < sl[s + 1]; p += 64) {
                const int d = r - cl[p];
                fits = fits && d >= 0 && d <= 0xffff;
                dl[p] = static_cast<unsigned short>(d);
            }
            for (int p = su[s] + l; p < su[s + 1]; p += 64) {
                const int d = cu[p] - r;
                fits = fits && d >= 0 && d <= 0xffff;
                du[p] = static_cast<unsigned short>(d);
            }
        }
    A->ell16 = false;
    FG_TRY(A->ell_sl.upload(sl.data(), sl.size(), ctx->stream));
    FG_TRY(A->ell_su.upload(su.data(), su.size(), ctx->stream));
    FG_TRY(A->ell_cl.upload(cl.data(), cl.size(), ctx->stream));
    FG_TRY(A->ell_cu.upload(cu.data(), cu.size(), ctx->stream));
    if (fits) {
        FG_TRY(A->ell_dl.upload(dl.data(), dl.size(), ctx->stream));
        FG_TRY(A->ell_du.upload(du.data(), du.size(), ctx->stream));
        A->ell16 = true;
    }
    // zero from allocation: the kernels that fill the values from parallel-ILU sweeps write only the
    // real entries, so the padding must already hold zeros
    FG_TRY(A->ell_vl.alloc(cl.size()));
    FG_TRY(A->ell_vu.alloc(cu.size()));
    FG_HIP(hipMemsetAsync(A->ell_vl.p, 0, sizeof(float) * cl.size(), ctx->stream));
    FG_HIP(hipMemsetAsync(A->ell_vu.p, 0, sizeof(float) * cu.size(), ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));   // the host arrays end with this call
    A->ell_kl = kl;
    A->ell_ku = ku;
    return FEMGPU_OK;
}

static int refresh_ell(femgpu_ctx *ctx, femgpu_csr *A)
{
    if (A->ell_kl == 0) return FEMGPU_OK;
    k_jacobi_ell_values<<<(A->n + 255) / 256, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->diag.p, A->LU32.p,
                                                                      A->ell_sl.p, A->ell_su.p, A->ell_vl.p,
                                                                      A->ell_vu.p);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

static int parilu_sell(femgpu_ctx *ctx, femgpu_csr *A, int sweeps)
{
    const int n = A->n, grid = xcd_grid((n + 255) / 256);
    // the final iterate must land in par_a / Dinv32
    float *vl = A->ell_kl > 0 ? A->ell_vl.p : nullptr, *vu = A->ell_kl > 0 ? A->ell_vu.p : nullptr;
    if (sweeps == 0) {   // the initial guess itself
        if (A->par_maxw <= 8)
            k_parilu_init_sell<float, 8><<<grid, 256, 0, ctx->stream>>>(n, A->sell_ptr.p, A->sell_col.p, A->sell_val.p,
                                                                 A->par_rinfo.p, A->par_dpos.p, A->par_a.p, A->Dinv32.p);
        else
            k_parilu_init_sell<float, 16><<<grid, 256, 0, ctx->stream>>>(n, A->sell_ptr.p, A->sell_col.p, A->sell_val.p,
                                                                  A->par_rinfo.p, A->par_dpos.p, A->par_a.p, A->Dinv32.p);
        if (vl)
            k_parilu_sell_to_ell<float><<<grid, 256, 0, ctx->stream>>>(n, A->sell_ptr.p, A->par_rinfo.p, A->par_a.p,
                                                                A->ell_sl.p, A->ell_su.p, vl, vu);
    } else if (A->par_maxw <= 8) {
        parilu_sweeps<float, 8>(ctx, A, sweeps, A->sell_val.p, A->par_a.p, A->par_b.p, A->Dinv32.p, A->Dinv32b.p, vl, vu);
    } else {
        parilu_sweeps<float, 16>(ctx, A, sweeps, A->sell_val.p, A->par_a.p, A->par_b.p, A->Dinv32.p, A->Dinv32b.p, vl,
                                 vu);
    }
    FG_HIP(hipGetLastError());
    A->lu32_in_sell = true;
    if (vl) return FEMGPU_OK;
    return launch_parilu_to_csr(ctx, A);   // the CSR Jacobi kernels read LU32
}

int launch_parilu32(femgpu_ctx *ctx, femgpu_csr *A, int sweeps)
{
    if (A->n == 0) return FEMGPU_OK;
    FG_TRY(ensure_par(A));
    if (A->par_ok) return parilu_sell(ctx, A, sweeps);
    A->lu32_in_sell = false;
    const FactorPlan &F = A->fplan_nat;
    const int grid = (A->n + 255) / 256;
    // the final sweep must land in LU32/Dinv32: start in whichever buffer makes the parity right
    float *L0 = (sweeps % 2 == 0) ? A->LU32.p : A->LU32b.p, *L1 = (sweeps % 2 == 0) ? A->LU32b.p : A->LU32.p;
    float *D0 = (sweeps % 2 == 0) ? A->Dinv32.p : A->Dinv32b.p, *D1 = (sweeps % 2 == 0) ? A->Dinv32b.p : A->Dinv32.p;
    k_parilu_init<<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->diag.p, A->AA32.p, L0, D0);
    FG_HIP(hipGetLastError());
    for (int s = 0; s < sweeps; ++s) {
        if (F.maxrow <= 8 && F.maxupd <= 16)
            k_parilu_sweep<8, 16><<<grid, 256, 0, ctx->stream>>>(A->n, F.rowinfo.p, F.upd.p, A->JA.p, A->AA32.p, L0, D0,
                                                                 L1, D1, nullptr);
        else
            k_parilu_sweep<1, 1><<<grid, 256, 0, ctx->stream>>>(A->n, F.rowinfo.p, F.upd.p, A->JA.p, A->AA32.p, L0, D0,
                                                                L1, D1, nullptr);
        FG_HIP(hipGetLastError());
        std::swap(L0, L1);
        std::swap(D0, D1);
    }
    return refresh_ell(ctx, A);
}

int launch_factor_to_f32(femgpu_ctx *ctx, femgpu_csr *A)
{
    FG_TRY(ensure_par(A));
    FG_TRY(launch_to_f32(ctx, A->LU.p, A->LU32.p, A->nnz));
    FG_TRY(launch_to_f32(ctx, A->Dinv.p, A->Dinv32.p, A->n));
    A->lu32_in_sell = false;
    return refresh_ell(ctx, A);
}

// x = (U~ L~)^-1 b approximated with it Jacobi iterations per triangle on factors in the Jacobi
// sliced-ELL layout (vl, vu strict triangles, dinv inverse pivots); ta, tb: two work vectors.
// b may alias x.
// rb != nullptr (FP32): b is not given but formed by the first lower launch as float(rb / *rn) and
// stored in bw (then b = bw for the later iterations).
template <typename T, int K, typename CT>
static void jacobi_ell_cols(femgpu_ctx *ctx, femgpu_csr *A, const CT *cl, const CT *cu, int it, const T *vl,
                            const T *vu, const T *dinv, T *ta, T *tb, const T *b, T *x, const int *skip,
                            const T *dv, double *dpart, const double *rb = nullptr, const double *rn = nullptr,
                            T *bw = nullptr)
{
    const int n = A->n, grid = xcd_grid((n + 256 * kJacRows - 1) / (256 * kJacRows));
    T *t[2] = {ta, tb};
    const T *prev = nullptr;
    for (int k = 0; k < it; ++k) {
        T *out = t[k & 1];
        if constexpr (std::is_same_v<T, float>) {
            if (k == 0 && rb) {
                k_jacobi_lower_ell_scaled<K, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_sl.p, cl, vl, rb, rn, bw, out);
                b = bw;
                prev = out;
                continue;
            }
        }
        k_jacobi_lower_ell<T, K, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_sl.p, cl, vl, b, prev, out, skip);
        prev = out;
    }
    const T *y = prev;
    T *other = (y == t[0]) ? t[1] : t[0];
    prev = nullptr;
    for (int k = 1; k <= it; ++k) {
        T *out = ((it - k) % 2 == 0) ? x : other;
        k_jacobi_upper_ell<T, K, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_su.p, cu, vu, dinv, y, prev, out, skip,
                                                                   dv, k == it ? dpart : nullptr);
        prev = out;
    }
}

template <typename T, int K>
static void jacobi_ell(femgpu_ctx *ctx, femgpu_csr *A, int it, const T *vl, const T *vu, const T *dinv, T *ta, T *tb,
                       const T *b, T *x, const int *skip, const T *dv = nullptr, double *dpart = nullptr,
                       const double *rb = nullptr, const double *rn = nullptr, T *bw = nullptr)
{
    if (A->ell16)
        jacobi_ell_cols<T, K, unsigned short>(ctx, A, A->ell_dl.p, A->ell_du.p, it, vl, vu, dinv, ta, tb, b, x, skip,
                                              dv, dpart, rb, rn, bw);
    else
        jacobi_ell_cols<T, K, int>(ctx, A, A->ell_cl.p, A->ell_cu.p, it, vl, vu, dinv, ta, tb, b, x, skip, dv, dpart,
                                   rb, rn, bw);
}

// The current values of A into the sliced-ELL layout in FP64 (padding 0).
__global__ void k_sell_fill64(int n, const int *__restrict__ IA, const int *__restrict__ sptr,
                              const double *__restrict__ AA, double *__restrict__ out)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int s = row >> 6, l = row & 63, p0 = sptr[s], w = (sptr[s + 1] - p0) >> 6;
    const int a0 = IA[row], len = IA[row + 1] - a0;
    for (int k = 0; k < w; ++k) out[p0 + (k << 6) + l] = k < len ? AA[a0 + k] : 0.0;
}

int launch_parilu64(femgpu_ctx *ctx, femgpu_csr *A, int sweeps)
{
    const int n = A->n;
    A->par64_factored = false;
    if (n == 0) {
        A->par64_factored = true;
        return FEMGPU_OK;
    }
    FG_CHECK(A->par_ok && A->ell_kl > 0, FEMGPU_ERR_ARG,
             "FP64 parallel ILU(0): needs the sliced-ELL plans (rows and strict triangles of at most 16 entries)");
    FG_CHECK(sweeps >= 1 && sweeps <= 1024, FEMGPU_ERR_ARG, "FP64 parallel ILU(0): %d sweeps out of range", sweeps);
    const size_t tot = A->sell_col.n, n1 = static_cast<size_t>(n) + 1;
    FG_TRY(A->par64_v.alloc(tot));
    FG_TRY(A->par64_a.alloc(tot));
    FG_TRY(A->par64_b.alloc(tot));
    FG_TRY(A->par64_d.alloc(n1));
    FG_TRY(A->par64_db.alloc(n1));
    if (A->ell64_vl.n < A->ell_cl.n || A->ell64_vu.n < A->ell_cu.n) {   // new arrays: padding zeros
        FG_TRY(A->ell64_vl.alloc(A->ell_cl.n));
        FG_TRY(A->ell64_vu.alloc(A->ell_cu.n));
        FG_HIP(hipMemsetAsync(A->ell64_vl.p, 0, sizeof(double) * A->ell64_vl.n, ctx->stream));
        FG_HIP(hipMemsetAsync(A->ell64_vu.p, 0, sizeof(double) * A->ell64_vu.n, ctx->stream));
    }
    FG_TRY(A->jac64_a.alloc(n1));
    FG_TRY(A->jac64_b.alloc(n1));
    k_sell_fill64<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, A->IA.p, A->sell_ptr.p, A->AA.p, A->par64_v.p);
    // the final iterate lands in par64_a / par64_d, its strict triangles in the Jacobi arrays
    if (A->par_maxw <= 8)
        parilu_sweeps<double, 8>(ctx, A, sweeps, A->par64_v.p, A->par64_a.p, A->par64_b.p, A->par64_d.p, A->par64_db.p,
                                 A->ell64_vl.p, A->ell64_vu.p);
    else
        parilu_sweeps<double, 16>(ctx, A, sweeps, A->par64_v.p, A->par64_a.p, A->par64_b.p, A->par64_d.p,
                                  A->par64_db.p, A->ell64_vl.p, A->ell64_vu.p);
    FG_HIP(hipGetLastError());
    A->par64_factored = true;
    return FEMGPU_OK;
}

int launch_jacobi_apply64(femgpu_ctx *ctx, femgpu_csr *A, int it, const double *b, double *x, const int *skip)
{
    const int n = A->n;
    if (n == 0) return FEMGPU_OK;
    FG_CHECK(it >= 1 && it <= 1024 && A->ell_kl > 0, FEMGPU_ERR_ARG, "FP64 Jacobi sweeps: %d iterations", it);
    const int w = std::max(A->ell_kl, A->ell_ku);
    const double *vl = A->ell64_vl.p, *vu = A->ell64_vu.p, *d = A->par64_d.p;
    double *ta = A->jac64_a.p, *tb = A->jac64_b.p;
    if (w <= 4) jacobi_ell<double, 4>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip);
    else if (w <= 8) jacobi_ell<double, 8>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip);
    else jacobi_ell<double, 16>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

template <typename CT>
static void jacobi_iteration_ell(femgpu_ctx *ctx, femgpu_csr *A, const CT *cl, const CT *cu, int w, int grid, int upper,
                                 const float *rhs, const float *in, float *out)
{
    const int n = A->n;
    if (upper) {
        if (w <= 4)
            k_jacobi_upper_ell<float, 4, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_su.p, cu, A->ell_vu.p, A->Dinv32.p,
                                                                         rhs, in, out, nullptr, nullptr, nullptr);
        else if (w <= 8)
            k_jacobi_upper_ell<float, 8, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_su.p, cu, A->ell_vu.p, A->Dinv32.p,
                                                                         rhs, in, out, nullptr, nullptr, nullptr);
        else
            k_jacobi_upper_ell<float, 16, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_su.p, cu, A->ell_vu.p, A->Dinv32.p,
                                                                          rhs, in, out, nullptr, nullptr, nullptr);
    } else {
        if (w <= 4)
            k_jacobi_lower_ell<float, 4, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_sl.p, cl, A->ell_vl.p, rhs, in, out,
                                                                         nullptr);
        else if (w <= 8)
            k_jacobi_lower_ell<float, 8, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_sl.p, cl, A->ell_vl.p, rhs, in, out,
                                                                         nullptr);
        else
            k_jacobi_lower_ell<float, 16, CT><<<grid, 256, 0, ctx->stream>>>(n, A->ell_sl.p, cl, A->ell_vl.p, rhs, in, out,
                                                                          nullptr);
    }
}

int launch_jacobi_apply32_scaled(femgpu_ctx *ctx, femgpu_csr *A, int it, const double *r, const double *rn, float *x,
                                 const float *dv, double *dpart)
{
    const int n = A->n;
    if (n == 0) return FEMGPU_OK;
    FG_CHECK(it >= 1 && it <= 1024 && A->ell_kl > 0, FEMGPU_ERR_ARG,
             "Jacobi sweeps from the FP64 residual: needs the sliced-ELL triangles (%d iterations)", it);
    const int w = std::max(A->ell_kl, A->ell_ku);
    const float *vl = A->ell_vl.p, *vu = A->ell_vu.p, *d = A->Dinv32.p;
    float *ta = A->jac_a.p, *tb = A->jac_b.p;
    if (w <= 4) jacobi_ell<float, 4>(ctx, A, it, vl, vu, d, ta, tb, x, x, nullptr, dv, dpart, r, rn, x);
    else if (w <= 8) jacobi_ell<float, 8>(ctx, A, it, vl, vu, d, ta, tb, x, x, nullptr, dv, dpart, r, rn, x);
    else jacobi_ell<float, 16>(ctx, A, it, vl, vu, d, ta, tb, x, x, nullptr, dv, dpart, r, rn, x);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int launch_jacobi_iteration32(femgpu_ctx *ctx, femgpu_csr *A, int upper, const float *rhs, const float *in,
                              float *out)
{
    const int n = A->n;
    if (n == 0) return FEMGPU_OK;
    const int grid = xcd_grid((n + 256 * kJacRows - 1) / (256 * kJacRows));
    if (A->ell_kl > 0) {
        const int w = std::max(A->ell_kl, A->ell_ku);
        if (A->ell16)
            jacobi_iteration_ell<unsigned short>(ctx, A, A->ell_dl.p, A->ell_du.p, w, grid, upper, rhs, in, out);
        else
            jacobi_iteration_ell<int>(ctx, A, A->ell_cl.p, A->ell_cu.p, w, grid, upper, rhs, in, out);
    } else {
        constexpr int LPR = 4;
        const int g4 = static_cast<int>((static_cast<long>(n) * LPR + 255) / 256);
        if (upper)
            k_jacobi_upper<LPR><<<g4, 256, 0, ctx->stream>>>(n, A->IA.p, A->JA.p, A->diag.p, A->LU32.p, A->Dinv32.p, rhs,
                                                             in, out, nullptr);
        else
            k_jacobi_lower<LPR><<<g4, 256, 0, ctx->stream>>>(n, A->IA.p, A->JA.p, A->diag.p, A->LU32.p, rhs, in, out,
                                                             nullptr);
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int jacobi_dot_blocks(const femgpu_csr *A)
{
    return A->ell_kl > 0 && A->n > 0 ? xcd_grid((A->n + 256 * kJacRows - 1) / (256 * kJacRows)) : 0;
}

int launch_jacobi_apply32(femgpu_ctx *ctx, femgpu_csr *A, int it, const float *b, float *x, const int *skip,
                          const float *dv, double *dpart)
{
    const int n = A->n;
    if (n == 0) return FEMGPU_OK;
    FG_CHECK(it >= 1 && it <= 1024, FEMGPU_ERR_ARG, "Jacobi sweeps: %d iterations out of range", it);
    FG_CHECK(!dpart || jacobi_dot_blocks(A) > 0, FEMGPU_ERR_ARG, "Jacobi sweeps: no fused product here");
    if (A->ell_kl > 0) {
        const int w = std::max(A->ell_kl, A->ell_ku);
        const float *vl = A->ell_vl.p, *vu = A->ell_vu.p, *d = A->Dinv32.p;
        float *ta = A->jac_a.p, *tb = A->jac_b.p;
        if (w <= 4) jacobi_ell<float, 4>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip, dv, dpart);
        else if (w <= 8) jacobi_ell<float, 8>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip, dv, dpart);
        else jacobi_ell<float, 16>(ctx, A, it, vl, vu, d, ta, tb, b, x, skip, dv, dpart);
        FG_HIP(hipGetLastError());
        return FEMGPU_OK;
    }
    constexpr int LPR = 4;
    const int grid = static_cast<int>((static_cast<long>(n) * LPR + 255) / 256);
    float *t[2] = {A->jac_a.p, A->jac_b.p};
    // lower: y_1 .. y_it alternate t[0], t[1]
    const float *prev = nullptr;
    for (int k = 0; k < it; ++k) {
        float *out = t[k & 1];
        k_jacobi_lower<LPR><<<grid, 256, 0, ctx->stream>>>(n, A->IA.p, A->JA.p, A->diag.p, A->LU32.p, b, prev, out, skip);
        prev = out;
    }
    const float *y = prev;
    float *other = (y == t[0]) ? t[1] : t[0];
    // upper: the last iterate goes to x; earlier ones alternate between x and the free buffer
    prev = nullptr;
    for (int k = 1; k <= it; ++k) {
        float *out = ((it - k) % 2 == 0) ? x : other;
        k_jacobi_upper<LPR><<<grid, 256, 0, ctx->stream>>>(n, A->IA.p, A->JA.p, A->diag.p, A->LU32.p, A->Dinv32.p, y,
                                                           prev, out, skip);
        prev = out;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

}  // namespace fg
