// Sparse kernels of the hot path for gfx950: CSR SpMV (FP64/FP32), residual, the chunked
// level-scheduled ILU(0) sweeps and numeric factorisation, the coloured EBE matvec, and the
// reduction / vector kernels. All memory-bound: no MFMA (3x3..12x12 blocks and CSR rows are not
// dense contractions).
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

// ------------------------------------------------------------------------------- CSR SpMV
// LPR lanes per row (a 64-wide wave covers 64/LPR consecutive rows, so the AA/JA reads of a wave
// are one contiguous span); partial row sums combined with xor-shuffles inside the lane group.
// The x operand is read as x[col]*xs with xs = *xs_ptr (1 when null): with xs = 1/h this fuses the
// Arnoldi normalisation of pgmres.c:113 into the product, and when xout is set lane 0 also stores
// the normalised basis entry xout[row] = x[row]*xs. A non-zero *skip makes the launch a no-op
// (used by the solvers to drain steps enqueued after convergence).
template <typename T, int LPR>
__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ IA, const int *__restrict__ JA,
                                              const T *__restrict__ AA, const T *__restrict__ x,
                                              const double *__restrict__ xs_ptr, T *__restrict__ xout,
                                              T *__restrict__ y, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int gt = xcd_block() * 256 + threadIdx.x;
    const int row = gt / LPR, lane = gt % LPR;
    if (row >= n) return;
    const T xs = xs_ptr ? static_cast<T>(*xs_ptr) : T(1);
    const int p0 = IA[row], p1 = IA[row + 1];
    T acc = 0;
    // two of the lane's entries per step, loaded together (the second clamped to the first when
    // past the row's end and then not added): the plain loop's additions, in its order
    for (int p = p0 + lane; p < p1; p += 2 * LPR) {
        const int p2 = p + LPR, q2 = p2 < p1 ? p2 : p;
        const T t1 = AA[p] * (x[JA[p]] * xs), t2 = AA[q2] * (x[JA[q2]] * xs);
        acc += t1;
        acc = p2 < p1 ? acc + t2 : acc;
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (lane == 0) {
        y[row] = acc;
        if (xout) xout[row] = x[row] * xs;
    }
}

// SpMV on the sliced-ELL copy: one row per thread, a wave's 64 rows form one slice padded to its
// longest row, entries column-major inside the slice, so every load of a step of the row loop is
// one contiguous run and the loop count is uniform across the wave. Same xs / xout / skip contract
// as k_spmv.
//
// k_spmv_sell<T, K> (slices at most K wide) issues all K column and value loads of the row before
// the first gather (steps past the slice width re-read step 0 and are not added, so the sum is the
// plain loop's over the row, in its order). Plain loads: non-temporal ones were faster in isolation
// but slower inside the solver, where the matrix is re-read every Arnoldi step and stays cached.
// k_spmv_sell_loop takes any width.
constexpr int kSell64MaxWidth = 16;   // widest slice the FP64 sliced-ELL kernels take

// Column of a sliced-ELL entry of row r: int columns as stored, 16-bit ones as the distance from r
// (2 bytes per entry instead of 4 when the band allows).
__device__ __forceinline__ int sell_colof(int c, int) { return c; }
__device__ __forceinline__ int sell_colof(short c, int r) { return r + static_cast<int>(c); }

template <typename T, int K, typename CT>
__global__ __launch_bounds__(256) void k_spmv_sell(int n, const int *__restrict__ sptr, const CT *__restrict__ col,
                                                   const T *__restrict__ val, const T *__restrict__ x,
                                                   const double *__restrict__ xs_ptr, T *__restrict__ xout,
                                                   T *__restrict__ y, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int row = xcd_block() * 256 + threadIdx.x;
    const int rr = row < n ? row : n - 1;
    const int s = rr >> 6, l = rr & 63;
    const int p0 = sptr[s], w = (sptr[s + 1] - p0) >> 6;
    int c[K];
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int e = p0 + ((k < w ? k : 0) << 6) + l;
        c[k] = sell_colof(col[e], rr);
        v[k] = val[e];
    }
    const T xs = xs_ptr ? static_cast<T>(*xs_ptr) : T(1);
    T acc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const T t = v[k] * (x[c[k]] * xs);
        acc = k < w ? acc + t : acc;
    }
    if (row >= n) return;
    y[row] = acc;
    if (xout) xout[row] = x[row] * xs;
}

template <typename T>
__global__ __launch_bounds__(256) void k_spmv_sell_loop(int n, const int *__restrict__ sptr,
                                                        const int *__restrict__ col, const T *__restrict__ val,
                                                        const T *__restrict__ x, const double *__restrict__ xs_ptr,
                                                        T *__restrict__ xout, T *__restrict__ y,
                                                        const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int row = xcd_block() * 256 + threadIdx.x;
    if (row >= n) return;
    const T xs = xs_ptr ? static_cast<T>(*xs_ptr) : T(1);
    const int s = row >> 6, l = row & 63;
    const int p0 = sptr[s], w = (sptr[s + 1] - p0) >> 6;
    T acc = 0;
    for (int k = 0; k < w; ++k) {
        const int e = p0 + (k << 6) + l;
        acc += val[e] * (x[col[e]] * xs);
    }
    y[row] = acc;
    if (xout) xout[row] = x[row] * xs;
}

// r = f - A x on the FP64 sliced-ELL copy (loads as in k_spmv_sell<double, K>).
template <int K, typename CT>
__global__ __launch_bounds__(256) void k_residual_sell(int n, const int *__restrict__ sptr, const CT *__restrict__ col,
                                                       const double *__restrict__ val, const double *__restrict__ x,
                                                       const double *__restrict__ f, double *__restrict__ r,
                                                       double *__restrict__ sumsq)
{
    __shared__ double red[4];
    const int blk = xcd_block(), row = blk * 256 + threadIdx.x;
    const int rr = row < n ? row : n - 1;
    const int s = rr >> 6, l = rr & 63;
    const int p0 = sptr[s], w = (sptr[s + 1] - p0) >> 6;
    int c[K];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int e = p0 + ((k < w ? k : 0) << 6) + l;
        c[k] = sell_colof(col[e], rr);
        v[k] = val[e];
    }
    const double fr = f[rr];
    double acc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = v[k] * x[c[k]];
        acc = k < w ? acc + t : acc;
    }
    const double rv = fr - acc;
    if (row < n) r[row] = rv;
    if (sumsq) {   // the block's partial of ||r||^2 (a later kernel sums the blocks in order)
        const double q = block_sum<256>(row < n ? rv * rv : 0.0, red);
        if (threadIdx.x == 0) sumsq[blk] = q;
    }
}

// The CSR values into the sliced-ELL layout (FP32, and FP64 when val64 is set). K > 0: slices at
// most K wide, the row's K values loaded together (positions past the row clamped to its first,
// or to the dummy slot nnz for an empty row); K = 0: any width.
template <int K>
__global__ __launch_bounds__(256) void k_sell_values(int n, const int *__restrict__ IA, const int *__restrict__ sptr,
                                                     const double *__restrict__ AA, float *__restrict__ val,
                                                     double *__restrict__ val64)
{
    const int row = xcd_block() * 256 + threadIdx.x;
    if (row >= n) return;
    const int s = row >> 6, l = row & 63;
    const int p0 = sptr[s], w = (sptr[s + 1] - p0) >> 6;
    const int a0 = IA[row], len = IA[row + 1] - a0;
    if constexpr (K > 0) {
        double a[K];
#pragma unroll
        for (int k = 0; k < K; ++k) a[k] = AA[a0 + (k < len ? k : 0)];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < w) {
                const double v = k < len ? a[k] : 0.0;
                val[p0 + (k << 6) + l] = static_cast<float>(v);
                if (val64) val64[p0 + (k << 6) + l] = v;
            }
    } else {
        for (int k = 0; k < w; ++k) {
            const double v = k < len ? AA[a0 + k] : 0.0;
            val[p0 + (k << 6) + l] = static_cast<float>(v);
            if (val64) val64[p0 + (k << 6) + l] = v;
        }
    }
}

int build_sell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA)
{
    const int n = A->n;
    A->sell_ok = 0;
    if (n == 0) return FEMGPU_OK;
    const int ns = (n + 63) / 64;
    std::vector<int> sptr(static_cast<size_t>(ns) + 1, 0);
    long long total = 0;
    int wmax = 0;
    for (int s = 0; s < ns; ++s) {
        int w = 0;
        for (int r = 64 * s; r < std::min(n, 64 * s + 64); ++r) w = std::max(w, IA[r + 1] - IA[r]);
        wmax = std::max(wmax, w);
        total += 64LL * w;
        if (total >= INT_MAX) return FEMGPU_OK;   // int entry offsets: stay with CSR
        sptr[static_cast<size_t>(s) + 1] = static_cast<int>(total);
    }
    if (total > 2LL * IA[n] + 64LL * ns) return FEMGPU_OK;   // rows too uneven: padding would dominate
    std::vector<int> col(static_cast<size_t>(total));
    for (int s = 0; s < ns; ++s) {
        const int w = (sptr[s + 1] - sptr[s]) / 64;
        for (int l = 0; l < 64; ++l) {
            const int r = 64 * s + l;
            for (int k = 0; k < w; ++k) {
                const bool has = r < n && k < IA[r + 1] - IA[r];
                col[static_cast<size_t>(sptr[s]) + 64 * k + l] = has ? JA[IA[r] + k] : std::min(r, n - 1);
            }
        }
    }
    // 16-bit column distances when the band allows
    std::vector<short> d16(col.size());
    bool fits = true;
    for (int s = 0; s < ns && fits; ++s)
        for (int l = 0; l < 64 && fits; ++l) {
            const int r = std::min(64 * s + l, n - 1);
            for (int p = sptr[s] + l; p < sptr[s + 1]; p += 64) {
                const int d = col[p] - r;
                fits = fits && d >= -32768 && d <= 32767;
                d16[p] = static_cast<short>(d);
            }
        }
    A->sell16 = false;
    A->hsell_ptr = sptr;
    FG_TRY(A->sell_ptr.upload(sptr.data(), sptr.size(), ctx->stream));
    FG_TRY(A->sell_col.upload(col.data(), col.size(), ctx->stream));
    if (fits) {
        FG_TRY(A->sell_d16.upload(d16.data(), d16.size(), ctx->stream));
        A->sell16 = true;
    }
    FG_TRY(A->sell_val.alloc(col.size()));
    if (A->spmv_lanes == 0 && wmax <= kSell64MaxWidth) FG_TRY(A->sell_val64.alloc(col.size()));
    FG_HIP(hipStreamSynchronize(ctx->stream));   // the host arrays end with this call
    A->sell_wmax = wmax;
    A->sell_ok = 1;
    return FEMGPU_OK;
}

int launch_sell_values(femgpu_ctx *ctx, femgpu_csr *A)
{
    if (!A->sell_ok) return FEMGPU_OK;
    const int grid = xcd_grid((A->n + 255) / 256);
    double *v64 = A->spmv_lanes == 0 ? A->sell_val64.p : nullptr;
    if (A->sell_wmax <= 8)
        k_sell_values<8><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->sell_ptr.p, A->AA.p, A->sell_val.p, v64);
    else if (A->sell_wmax <= 16)
        k_sell_values<16><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->sell_ptr.p, A->AA.p, A->sell_val.p, v64);
    else
        k_sell_values<0><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->sell_ptr.p, A->AA.p, A->sell_val.p, v64);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

bool sell64_possible(const femgpu_csr *A) { return A->sell_ok && A->sell_wmax <= kSell64MaxWidth; }

bool sell64_active(const femgpu_csr *A) { return A->spmv_lanes == 0 && sell64_possible(A) && A->sell_val64.p; }

int set_spmv_lanes(femgpu_csr *A, int lanes)
{
    if (lanes == 0) {
        if (!sell64_possible(A)) {
            set_error("FP64 sliced ELL needs the sliced-ELL copy with slices at most %d wide", kSell64MaxWidth);
            return FEMGPU_ERR_ARG;
        }
        const bool fresh = A->spmv_lanes != 0;
        FG_TRY(A->sell_val64.alloc(A->sell_col.n));
        A->spmv_lanes = 0;
        if (fresh) FG_TRY(launch_sell_values(A->ctx, A));   // fill from the current values
        return FEMGPU_OK;
    }
    A->spmv_lanes = lanes;
    A->sell_val64.release();
    return FEMGPU_OK;
}

// r = f - A x (FP64 outer residual of the mixed solver).
template <int LPR>
__global__ __launch_bounds__(256) void k_residual(int n, const int *__restrict__ IA, const int *__restrict__ JA,
                                                  const double *__restrict__ AA, const double *__restrict__ x,
                                                  const double *__restrict__ f, double *__restrict__ r)
{
    const int gt = xcd_block() * 256 + threadIdx.x;
    const int row = gt / LPR, lane = gt % LPR;
    if (row >= n) return;
    const int p0 = IA[row], p1 = IA[row + 1];
    double acc = 0;
    for (int p = p0 + lane; p < p1; p += 2 * LPR) {   // as in k_spmv: two entries per step
        const int p2 = p + LPR, q2 = p2 < p1 ? p2 : p;
        const double t1 = AA[p] * x[JA[p]], t2 = AA[q2] * x[JA[q2]];
        acc += t1;
        acc = p2 < p1 ? acc + t2 : acc;
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (lane == 0) r[row] = f[row] - acc;
}

template <typename T>
static int spmv_dispatch(femgpu_ctx *ctx, const femgpu_csr *A, const T *AA, const T *x, const double *xs,
                         T *xout, T *y, const int *skip)
{
    const int L = A->spmv_lanes > 0 ? A->spmv_lanes : 8;   // 0 (sliced ELL) is dispatched before this
    const long threads = static_cast<long>(A->n) * L;
    const int grid = xcd_grid(static_cast<int>((threads + 255) / 256));
    if (A->n == 0) return FEMGPU_OK;
    switch (L) {
    case 1: k_spmv<T, 1><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, AA, x, xs, xout, y, skip); break;
    case 2: k_spmv<T, 2><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, AA, x, xs, xout, y, skip); break;
    case 4: k_spmv<T, 4><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, AA, x, xs, xout, y, skip); break;
    case 16: k_spmv<T, 16><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, AA, x, xs, xout, y, skip); break;
    default: k_spmv<T, 8><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, AA, x, xs, xout, y, skip); break;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

template <typename T>
static int sell_dispatch(femgpu_ctx *ctx, const femgpu_csr *A, const T *val, const T *x, const double *xs, T *xout,
                         T *y, const int *skip)
{
    const int n = A->n, grid = xcd_grid((n + 255) / 256);
    const int *sp = A->sell_ptr.p, *sc = A->sell_col.p;
    const short *sd = A->sell_d16.p;
    if (A->sell_wmax <= 8) {
        if (A->sell16)
            k_spmv_sell<T, 8, short><<<grid, 256, 0, ctx->stream>>>(n, sp, sd, val, x, xs, xout, y, skip);
        else
            k_spmv_sell<T, 8, int><<<grid, 256, 0, ctx->stream>>>(n, sp, sc, val, x, xs, xout, y, skip);
    } else if (A->sell_wmax <= 16) {
        if (A->sell16)
            k_spmv_sell<T, 16, short><<<grid, 256, 0, ctx->stream>>>(n, sp, sd, val, x, xs, xout, y, skip);
        else
            k_spmv_sell<T, 16, int><<<grid, 256, 0, ctx->stream>>>(n, sp, sc, val, x, xs, xout, y, skip);
    } else {
        k_spmv_sell_loop<T><<<grid, 256, 0, ctx->stream>>>(n, sp, sc, val, x, xs, xout, y, skip);
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int launch_spmv64(femgpu_ctx *ctx, const femgpu_csr *A, const double *x, const double *xs, double *xout,
                  double *y, const int *skip)
{
    if (A->n > 0 && sell64_active(A)) return sell_dispatch<double>(ctx, A, A->sell_val64.p, x, xs, xout, y, skip);
    return spmv_dispatch<double>(ctx, A, A->AA.p, x, xs, xout, y, skip);
}

int launch_spmv32(femgpu_ctx *ctx, const femgpu_csr *A, const float *x, const double *xs, float *xout,
                  float *y, const int *skip)
{
    if (A->sell_ok && A->n > 0) return sell_dispatch<float>(ctx, A, A->sell_val.p, x, xs, xout, y, skip);
    return spmv_dispatch<float>(ctx, A, A->AA32.p, x, xs, xout, y, skip);
}

// blocks of the fused residual launch (the XCD-rounded grid; padding blocks write zero partials)
int residual_sumsq_blocks(const femgpu_csr *A) { return sell64_active(A) && A->n > 0 ? xcd_grid((A->n + 255) / 256) : 0; }

int launch_residual64(femgpu_ctx *ctx, const femgpu_csr *A, const double *x, const double *f, double *r,
                      double *sumsq)
{
    if (A->n == 0) return FEMGPU_OK;
    FG_CHECK(!sumsq || residual_sumsq_blocks(A) > 0, FEMGPU_ERR_ARG, "residual: no fused sum of squares here");
    if (sell64_active(A)) {
        const int n = A->n, grid = xcd_grid((n + 255) / 256);
        const int *sp = A->sell_ptr.p, *sc = A->sell_col.p;
        const short *sd = A->sell_d16.p;
        const double *sv = A->sell_val64.p;
        if (A->sell_wmax <= 8) {   // sell64_active implies sell_wmax <= kSell64MaxWidth
            if (A->sell16)
                k_residual_sell<8, short><<<grid, 256, 0, ctx->stream>>>(n, sp, sd, sv, x, f, r, sumsq);
            else
                k_residual_sell<8, int><<<grid, 256, 0, ctx->stream>>>(n, sp, sc, sv, x, f, r, sumsq);
        } else {
            if (A->sell16)
                k_residual_sell<kSell64MaxWidth, short><<<grid, 256, 0, ctx->stream>>>(n, sp, sd, sv, x, f, r, sumsq);
            else
                k_residual_sell<kSell64MaxWidth, int><<<grid, 256, 0, ctx->stream>>>(n, sp, sc, sv, x, f, r, sumsq);
        }
        FG_HIP(hipGetLastError());
        return FEMGPU_OK;
    }
    const int L = A->spmv_lanes > 0 ? A->spmv_lanes : 8;
    const int grid = xcd_grid(static_cast<int>((static_cast<long>(A->n) * L + 255) / 256));
    switch (L) {
    case 1: k_residual<1><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->AA.p, x, f, r); break;
    case 2: k_residual<2><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->AA.p, x, f, r); break;
    case 4: k_residual<4><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->AA.p, x, f, r); break;
    case 16: k_residual<16><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->AA.p, x, f, r); break;
    default: k_residual<8><<<grid, 256, 0, ctx->stream>>>(A->n, A->IA.p, A->JA.p, A->AA.p, x, f, r); break;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

__global__ void k_fill(int n, double v, double *__restrict__ x)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) x[k] = v;
}

// Rows-per-wavefront autotuning of the FP64 CSR kernels (SpMV and residual share the choice): each
// candidate lanes-per-row L (64 / L rows per wavefront) is timed with HIP events over 20 launches on
// the matrix's own structure and values, after two warm-up launches; the fastest is kept.
int csr_autotune(femgpu_csr *A, femgpu_tune_info *info)
{
    femgpu_ctx *ctx = A->ctx;
    // candidates: CSR with 1, 2, 4, 8, 16 lanes per row, then FP64 sliced ELL (lanes 0) where possible
    static constexpr int kLanes[6] = {1, 2, 4, 8, 16, 0};
    femgpu_tune_info out{};
    out.lanes = A->spmv_lanes;
    if (A->n == 0) {
        if (info) *info = out;
        return FEMGPU_OK;
    }
    DevBuf<double> x, y;
    FG_TRY(x.alloc(A->n));
    FG_TRY(y.alloc(A->n));
    k_fill<<<std::min(4096, (A->n + 255) / 256), 256, 0, ctx->stream>>>(A->n, 1.0, x.p);
    FG_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    FG_HIP(hipEventCreate(&e0));
    FG_HIP(hipEventCreate(&e1));
    const int saved = A->spmv_lanes;
    const int ncand = sell64_possible(A) ? 6 : 5;
    int best = saved, rc = FEMGPU_OK;
    float best_ms = 0.0f;
    for (int c = 0; c < ncand && rc == FEMGPU_OK; ++c) {
        if (kLanes[c] == 0) rc = set_spmv_lanes(A, 0);   // allocates and fills the FP64 sliced-ELL values
        else A->spmv_lanes = kLanes[c];
        for (int w = 0; w < 2 && rc == FEMGPU_OK; ++w) rc = launch_spmv64(ctx, A, x.p, nullptr, nullptr, y.p, nullptr);
        if (rc == FEMGPU_OK && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = FEMGPU_ERR_HIP;
        for (int t = 0; t < 20 && rc == FEMGPU_OK; ++t) rc = launch_spmv64(ctx, A, x.p, nullptr, nullptr, y.p, nullptr);
        float ms = 0.0f;
        if (rc == FEMGPU_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = FEMGPU_ERR_HIP;
        if (rc != FEMGPU_OK) break;
        (kLanes[c] == 0 ? out.us_sell : out.us[c]) = 1e3 * ms / 20.0;
        if (c == 0 || ms < best_ms) {
            best_ms = ms;
            best = kLanes[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int keep = rc == FEMGPU_OK ? best : saved;
    A->spmv_lanes = 8;   // set_spmv_lanes(0) then (re)fills the FP64 sliced-ELL values
    const int rc2 = set_spmv_lanes(A, keep);
    out.lanes = A->spmv_lanes;
    if (info) *info = out;
    if (rc != FEMGPU_OK) set_error("femgpu_csr_autotune: timing launches failed");
    return rc != FEMGPU_OK ? rc : rc2;
}

// ------------------------------------------------------------------------------- conversions
__global__ void k_to_f32(const double *__restrict__ s, float *__restrict__ d, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = static_cast<float>(s[i]);
}

int launch_to_f32(femgpu_ctx *ctx, const double *src, float *dst, size_t n)
{
    if (!n) return FEMGPU_OK;
    const int grid = static_cast<int>(std::min<size_t>((n + 255) / 256, 4096));
    k_to_f32<<<grid, 256, 0, ctx->stream>>>(src, dst, n);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

// ------------------------------------------------------------------------------- ILU(0) factorisation
// Numeric ILU(0) by one persistent workgroup walking the level schedule of the lower sweep: rows of
// a level in parallel, a barrier between levels. Inside one workgroup the barrier alone makes the
// finished rows visible (one CU, one L1), so no agent-scope synchronisation is needed.
constexpr int kFactorThreads = 1024;

__global__ __launch_bounds__(kFactorThreads) void k_ilu0_factor_levels(int depth, const int *__restrict__ lvl_ptr,
                                                                       const int *__restrict__ row_of,
                                                                       const int *__restrict__ IA, const int *__restrict__ JA,
                                                                       const int *__restrict__ diag, double *LU, double *Dinv,
                                                                       int *err)
{
#pragma clang fp contract(off)
    for (int l = 0; l < depth; ++l) {
        const int q1 = lvl_ptr[l + 1];
        for (int q = lvl_ptr[l] + threadIdx.x; q < q1; q += kFactorThreads) {
            const int i = row_of[q];
            const int p0 = IA[i], d = diag[i], p1 = IA[i + 1];
            double dv = LU[d];
            for (int j = p0; j < d; ++j) {
                const int jrow = JA[j];
                const double lij = LU[j] * Dinv[jrow];
                LU[j] = lij;
                const int ke = IA[jrow + 1];
                for (int kk = diag[jrow] + 1; kk < ke; ++kk) {
                    const int col = JA[kk];
                    if (col == i) {
                        dv -= lij * LU[kk];
                        continue;
                    }
                    int lo = p0, hi = p1 - 1, pos = -1;
                    while (lo <= hi) {
                        const int m = (lo + hi) >> 1;
                        const int cm = JA[m];
                        if (cm == col) { pos = m; break; }
                        if (cm < col) lo = m + 1; else hi = m - 1;
                    }
                    if (pos >= 0) LU[pos] -= lij * LU[kk];
                }
            }
            LU[d] = dv;
            if (dv == 0.0) {
                __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                Dinv[i] = 0.0;
            } else {
                Dinv[i] = 1.0 / dv;
            }
        }
        __syncthreads();
    }
}

// Numeric ILU(0) from the update lists: one persistent workgroup, rows of a level in parallel (one
// thread per row), a barrier between levels. A row first issues all of its loads at once -- its
// update list, the finished U values and pivots it reads, its own values -- so a level costs a few
// memory round trips instead of one per update; the row's values and multipliers live in a private
// LDS slice. Same operations in the same order as the serial factorisation, contraction off.
// Longer update lists are taken in chunks of MAXU. Rows beyond the compile-time bounds (MAXROW
// values, MAXL multipliers) take a sequential path with the same arithmetic.
template <int MAXROW, int MAXL, int MAXU, int THREADS>
__global__ __launch_bounds__(THREADS) void k_ilu0_factor_lists(int depth, const int *__restrict__ lvl_ptr,
                                                               const int4 *__restrict__ rowinfo,
                                                               const int2 *__restrict__ upd, const int *__restrict__ JA,
                                                               const double *__restrict__ AA, double *LU, double *Dinv,
                                                               int *err)
{
#pragma clang fp contract(off)
    __shared__ double scratch[(MAXROW + MAXL) * THREADS];
    double *v = scratch + threadIdx.x;                 // row value e at v[e * THREADS]
    double *dj = scratch + MAXROW * THREADS + threadIdx.x;   // pivot inverse of L entry t at dj[t * THREADS]
    for (int l = 0; l < depth; ++l) {
        const int q1 = lvl_ptr[l + 1];
        for (int q = lvl_ptr[l] + threadIdx.x; q < q1; q += THREADS) {
            const int4 ri = rowinfo[q];
            const int p0 = ri.x, len = ri.y, nL = ri.z, ub = ri.w;
            const int nu = rowinfo[q + 1].w - ub;   // rowinfo has a sentinel entry after the last row
            const int i = JA[p0 + nL];
            double piv;
            if (len <= MAXROW && nL <= MAXL) {
                int jr[MAXL];
#pragma unroll
                for (int t = 0; t < MAXL; ++t)
                    if (t < nL) jr[t] = JA[p0 + t];
#pragma unroll
                for (int e = 0; e < MAXROW; ++e)
                    if (e < len) v[e * THREADS] = AA[p0 + e];
#pragma unroll
                for (int t = 0; t < MAXL; ++t)
                    if (t < nL) dj[t * THREADS] = Dinv[jr[t]];
                int sc = 0;   // L entries [0, sc) already turned into multipliers
                // the update list in chunks of MAXU: each chunk's list entries and U values are
                // loaded together, then applied in list order
                for (int u0 = 0; u0 < nu; u0 += MAXU) {
                    int2 U[MAXU];
                    double Uv[MAXU];
#pragma unroll
                    for (int u = 0; u < MAXU; ++u)
                        if (u0 + u < nu) U[u] = upd[ub + u0 + u];
#pragma unroll
                    for (int u = 0; u < MAXU; ++u)
                        if (u0 + u < nu) Uv[u] = LU[U[u].x];
#pragma unroll
                    for (int u = 0; u < MAXU; ++u) {
                        if (u0 + u < nu) {
                            const int t = U[u].y >> 16, dest = U[u].y & 0xffff;
                            for (; sc <= t; ++sc) v[sc * THREADS] = v[sc * THREADS] * dj[sc * THREADS];
                            v[dest * THREADS] -= v[t * THREADS] * Uv[u];
                        }
                    }
                }
                for (; sc < nL; ++sc) v[sc * THREADS] = v[sc * THREADS] * dj[sc * THREADS];
#pragma unroll
                for (int e = 0; e < MAXROW; ++e)
                    if (e < len) LU[p0 + e] = v[e * THREADS];
                piv = (nL < MAXROW) ? v[nL * THREADS] : LU[p0 + nL];
            } else {
                // sequential path directly on the factors in global memory
                for (int e = 0; e < len; ++e) LU[p0 + e] = AA[p0 + e];
                int u = ub;
                const int ue = ub + nu;
                for (int t = 0; t < nL; ++t) {
                    const double lij = LU[p0 + t] * Dinv[JA[p0 + t]];
                    LU[p0 + t] = lij;
                    for (; u < ue && (upd[u].y >> 16) == t; ++u)
                        LU[p0 + (upd[u].y & 0xffff)] -= lij * LU[upd[u].x];
                }
                piv = LU[p0 + nL];
            }
            if (piv == 0.0) {
                __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                Dinv[i] = 0.0;
            } else {
                Dinv[i] = 1.0 / piv;
            }
        }
        __syncthreads();
    }
}

int launch_ilu_factor(femgpu_ctx *ctx, femgpu_csr *A)
{
    if (A->n == 0) return FEMGPU_OK;
    const FactorPlan &F = A->fplan;
    if (A->sfac.ok && A->factor_kernel == 0) {
        FG_TRY(launch_stream_factor(ctx, A));
    } else if (F.maxrow <= 8) {
        k_ilu0_factor_lists<8, 8, 24, 512><<<1, 512, 0, ctx->stream>>>(A->plo.depth, A->plo.lvl_ptr.p, F.rowinfo.p,
                                                                        F.upd.p, A->JA.p, A->AA.p, A->LU.p, A->Dinv.p,
                                                                        ctx->err.p);
    } else if (F.maxrow <= 32) {
        // 256 threads: the levels of the 9x9-block systems hold up to ~260 rows (513 x 513 cavity:
        // 180 ms per factorisation against 220 ms with 128 threads); 96 KB of LDS
        k_ilu0_factor_lists<32, 16, 32, 256><<<1, 256, 0, ctx->stream>>>(A->plo.depth, A->plo.lvl_ptr.p, F.rowinfo.p,
                                                                          F.upd.p, A->JA.p, A->AA.p, A->LU.p, A->Dinv.p,
                                                                          ctx->err.p);
    } else {
        FG_HIP(hipMemcpyAsync(A->LU.p, A->AA.p, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, ctx->stream));
        k_ilu0_factor_levels<<<1, kFactorThreads, 0, ctx->stream>>>(A->plo.depth, A->plo.lvl_ptr.p, A->plo.row_of.p,
                                                                    A->IA.p, A->JA.p, A->diag.p, A->LU.p, A->Dinv.p,
                                                                    ctx->err.p);
    }
    FG_HIP(hipGetLastError());
    return packed_refresh(ctx, A);
}

// ------------------------------------------------------------------------------- dot product
// Two kernels: per-block partials, then one block sums them in a fixed order (deterministic, no
// atomics). The inputs may be any device vectors, so the loads are scalar.
__global__ __launch_bounds__(256) void k_dot(int n, const double *__restrict__ x, const double *__restrict__ y,
                                             double *partials)
{
    __shared__ double red[256 / kWave];
    double acc = 0;
    // 64-bit index: n may lie within one grid stride of INT_MAX
    for (long i = blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) acc += x[i] * y[i];
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_dot_finish(int nb, const double *__restrict__ partials, double *out)
{
    __shared__ double red[256 / kWave];
    double v = 0;
    for (int b = threadIdx.x; b < nb; b += 256) v += partials[b];
    v = block_sum<256>(v, red);
    if (threadIdx.x == 0) *out = v;
}

int dot64_into(femgpu_ctx *ctx, int n, const double *x, const double *y, double *partials, double *d_out)
{
    const int grid = std::max(1, std::min(kReduceBlocks, (n + 1023) / 1024));
    k_dot<<<grid, 256, 0, ctx->stream>>>(n, x, y, partials);
    k_dot_finish<<<1, 256, 0, ctx->stream>>>(grid, partials, d_out);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int dot64(femgpu_ctx *ctx, int n, const double *x, const double *y, double *d_out)
{
    return dot64_into(ctx, n, x, y, ctx->partials.p, d_out);
}

}  // namespace fg
