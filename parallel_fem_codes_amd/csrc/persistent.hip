// Persistent batched mixed-precision GMRES: one workgroup per system runs the whole solve of its
// system in one launch -- the FP64 residual and refinement, the FP32 inner GMRES(m) on the sliced-ELL
// copy of A with the Jacobi-applied parallel ILU(0) factors, fused MGS passes, and the Givens and
// least-squares steps in LDS. Every system has its own Krylov space and Hessenberg matrix, so a
// system's iterates do not depend on the other systems of the launch: its answer is bit-identical
// alone and in any batch. For the small systems of the serving loop, where one launch per kernel
// step would cost more than the step's memory traffic.
//
// Synchronisation is workgroup-local only (__syncthreads), every loop is bounded (max_refinements
// outer steps, m - 1 inner steps, the Jacobi iteration count), and every exit decision is taken
// from values that all threads of the workgroup hold identically, so every wave reaches the end.
#include <algorithm>

#include "common.h"
#include "device_util.h"

namespace fg {
namespace {

constexpr int kWgMaxBasis = 64;   // inner basis size limit: the Hessenberg matrix lives in LDS

struct WgArgs {
    const int *row_off;
    const int *IA, *JA;          // FP64 CSR (outer residual)
    const double *AA;
    const int *sptr, *scol;      // sliced ELL of A (FP32 SpMV) and of the factors
    const float *sval;
    const int *rinfo;            // row -> strict-lower count | row length << 8
    const float *LU, *Dinv;      // parallel ILU(0) factors in the sliced-ELL layout, 1/diag(U)
    const double *F;
    double *X, *R;
    float *V, *T;                // basis (m + 1 vectors) and three work vectors, ldv apart
    size_t ldv;
    int kin, jac, max_ref;
    double tol, inner_tol;
    femgpu_batch_info *info;
};

// Sum over the workgroup, the same value (same order) in every thread. Two barriers: red may be
// reused by the next call at once.
template <int BS>
__device__ __forceinline__ double wg_sum(double v, double *red)
{
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < BS / 64; ++k) t += red[k];
    return t;
}

// Row loops are unrolled so that a thread has several rows' loads in flight, and the entry loop of
// a row is unrolled to the matrix's row-width bound QW with its loads unconditional (positions past
// the row's end are clamped to its first entry and their terms not added): the loads of a row go
// out together instead of one dependent round trip per entry, and the sums are the same for any QW.
template <int QW>
__device__ __forceinline__ float sell_row_dot(const WgArgs &a, const float *vals, int p, int q0, int q1,
                                              const float *x)
{
    float t = 0.0f;
#pragma unroll
    for (int u = 0; u < QW; ++u) {
        const int q = q0 + u;
        const bool in = q < q1;
        const int e = p + ((in ? q : 0) << 6);
        const float v = vals[e] * x[a.scol[e]];
        t = in ? t + v : t;
    }
    return t;
}

// z = (U~ L~)^-1 b approximated by jac Jacobi iterations per triangle (the iteration of the
// multi-kernel path): y_k = b - L y_{k-1} (y_0 = b), then z_k = Dinv (y - U z_{k-1})
// (z_0 = Dinv y, formed first into a free buffer). b is T0; the lower iterates go to T1/T2, the
// upper ones alternate over the two free buffers, and the last one lands in out.
template <int BS, int QW>
__device__ void wg_jacobi(const WgArgs &a, int r0, int r1, float *T0, float *T1, float *T2, float *out)
{
    const float *prev = nullptr;
    for (int k = 0; k < a.jac; ++k) {
        float *dst = (k & 1) ? T2 : T1;
        const float *src = prev ? prev : T0;
#pragma unroll 2
        for (int r = r0 + threadIdx.x; r < r1; r += BS) {
            const int p = a.sptr[r >> 6] + (r & 63), nl = a.rinfo[r] & 0xff;
            dst[r] = T0[r] - sell_row_dot<QW>(a, a.LU, p, 0, nl, src);
        }
        __syncthreads();
        prev = dst;
    }
    const float *y = prev;
    float *fa = T0, *fb = (y == T1) ? T2 : T1;
    // z_0 = Dinv y into fb; the iterates then alternate fa, fb, ..., the last one into out
#pragma unroll 2
    for (int r = r0 + threadIdx.x; r < r1; r += BS) fb[r] = a.Dinv[r] * y[r];
    __syncthreads();
    const float *zin = fb;
    for (int k = 1; k <= a.jac; ++k) {
        float *dst = (k == a.jac) ? out : ((k & 1) ? fa : fb);
#pragma unroll 2
        for (int r = r0 + threadIdx.x; r < r1; r += BS) {
            const int p = a.sptr[r >> 6] + (r & 63), inf = a.rinfo[r];
            const int nl = inf & 0xff, len = inf >> 8;
            dst[r] = a.Dinv[r] * (y[r] - sell_row_dot<QW>(a, a.LU, p, nl + 1, len, zin));
        }
        __syncthreads();
        zin = dst;
    }
}

template <int BS, int QW>
__global__ __launch_bounds__(BS) void k_gmres_wg(WgArgs a)
{
    extern __shared__ double sm[];
    __shared__ int s_done;
    const int ld = a.kin + 1;
    double *H = sm, *cs = H + ld * ld, *sn = cs + ld, *e = sn + ld, *y = e + ld, *red = y + ld;
    const int sys = blockIdx.x, tid = threadIdx.x;
    const int r0 = a.row_off[sys], r1 = a.row_off[sys + 1];
    float *T0 = a.T, *T1 = a.T + a.ldv, *T2 = a.T + 2 * a.ldv;

    double acc = 0.0;
    for (int r = r0 + tid; r < r1; r += BS) {
        const double f = a.F[r];
        acc += f * f;
        a.X[r] = 0.0;
    }
    const double normF = sqrt(wg_sum<BS>(acc, red));
    int refinements = 0, inner = 0;
    double rn = normF;
    while (normF > 0.0) {
        // outer FP64 residual r = F - A X (X = 0 on the first pass), CSR entries in row order
        acc = 0.0;
#pragma unroll 2
        for (int r = r0 + tid; r < r1; r += BS) {
            double v = a.F[r];
            if (refinements > 0) {
                const int p0 = a.IA[r], len = a.IA[r + 1] - p0;
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < QW; ++q) {
                    const bool in = q < len;
                    const int p = p0 + (in ? q : 0);
                    const double t = a.AA[p] * a.X[a.JA[p]];
                    s = in ? s + t : s;
                }
                v = v - s;
            }
            a.R[r] = v;
            acc += v * v;
        }
        rn = sqrt(wg_sum<BS>(acc, red));
        if (rn <= a.tol * normF || refinements >= a.max_ref) break;
        ++refinements;
        // inner FP32 solve of M^-1 A d = M^-1 r / ||r||, started from d = 0
        const double inv_rn = 1.0 / rn;
#pragma unroll 4
        for (int r = r0 + tid; r < r1; r += BS) T0[r] = static_cast<float>(a.R[r] * inv_rn);
        __syncthreads();
        float *V0 = a.V;
        wg_jacobi<BS, QW>(a, r0, r1, T0, T1, T2, V0);
        acc = 0.0;
#pragma unroll 4
        for (int r = r0 + tid; r < r1; r += BS) acc += static_cast<double>(V0[r]) * V0[r];
        const double e0 = sqrt(wg_sum<BS>(acc, red));
        if (!(e0 > 0.0)) break;   // preconditioned residual vanished in FP32: nothing left to correct
        const double eps = a.inner_tol * e0;
        const float inv0 = static_cast<float>(1.0 / e0);
#pragma unroll 4
        for (int r = r0 + tid; r < r1; r += BS) V0[r] = V0[r] * inv0;
        if (tid == 0) e[0] = e0;
        __syncthreads();
        int steps = 0;
        for (int i = 0; i < a.kin - 1; ++i) {
            const float *Vi = a.V + i * a.ldv;
            float *W = a.V + (i + 1) * a.ldv;
#pragma unroll 2
            for (int r = r0 + tid; r < r1; r += BS) {   // T0 = A v_i (sliced ELL, FP32)
                const int p = a.sptr[r >> 6] + (r & 63), len = a.rinfo[r] >> 8;
                T0[r] = sell_row_dot<QW>(a, a.sval, p, 0, len, Vi);
            }
            __syncthreads();
            wg_jacobi<BS, QW>(a, r0, r1, T0, T1, T2, W);   // W = M^-1 A v_i
            // MGS: pass j subtracts h_{j-1,i} v_{j-1} and accumulates <v_j, w> (j = i + 1: ||w||^2)
            double hprev = 0.0;
            for (int j = 0; j <= i + 1; ++j) {
                const float *Vj = a.V + j * a.ldv, *Vp = a.V + (j > 0 ? j - 1 : 0) * a.ldv;
                const float hp = static_cast<float>(hprev);
                acc = 0.0;
#pragma unroll 4
                for (int r = r0 + tid; r < r1; r += BS) {
                    float wv = W[r];
                    if (j > 0) {
                        wv = wv - hp * Vp[r];
                        W[r] = wv;
                    }
                    const float u = j <= i ? Vj[r] : wv;
                    acc += static_cast<double>(u) * static_cast<double>(wv);
                }
                const double h = wg_sum<BS>(acc, red);
                if (tid == 0 && j > 0) H[(j - 1) * ld + i] = hprev;
                hprev = h;
            }
            const double hn = sqrt(hprev);
            if (tid == 0) {
#pragma clang fp contract(off)
                H[(i + 1) * ld + i] = hn;
                for (int j = 0; j <= i - 1; ++j) {   // pgmres.c:118-159 on this system's column
                    const double h1 = cs[j] * H[j * ld + i] + sn[j] * H[(j + 1) * ld + i];
                    const double h2 = -sn[j] * H[j * ld + i] + cs[j] * H[(j + 1) * ld + i];
                    H[j * ld + i] = h1;
                    H[(j + 1) * ld + i] = h2;
                }
                const double ha = H[i * ld + i], hb = H[(i + 1) * ld + i];
                const double rr = sqrt(ha * ha + hb * hb);
                cs[i] = ha / rr;
                sn[i] = hb / rr;
                H[i * ld + i] = rr;
                H[(i + 1) * ld + i] = 0.0;
                e[i + 1] = -sn[i] * e[i];
                e[i] = cs[i] * e[i];
                s_done = !(fabs(e[i + 1]) > eps) || (i + 1 >= a.kin - 1);
            }
            __syncthreads();
            steps = i + 1;
            if (s_done) break;
            const float invh = static_cast<float>(1.0 / hn);
#pragma unroll 4
            for (int r = r0 + tid; r < r1; r += BS) W[r] = W[r] * invh;
            __syncthreads();
        }
        inner += steps;
        if (tid == 0) {   // back substitution (pgmres.c:159-170)
#pragma clang fp contract(off)
            const int i = steps - 1;
            y[i] = e[i] / H[i * ld + i];
            for (int j = i - 1; j >= 0; --j) {
                double soma = 0.0;
                for (int k = j + 1; k <= i; ++k) soma = soma + H[j * ld + k] * y[k];
                y[j] = (e[j] - soma) / H[j * ld + j];
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int r = r0 + tid; r < r1; r += BS) {   // X += ||r|| * V y
            double d = 0.0;
            for (int j = 0; j < steps; ++j) d = d + static_cast<double>(a.V[j * a.ldv + r]) * y[j];
            a.X[r] = a.X[r] + rn * d;
        }
        __syncthreads();
    }
    if (tid == 0) {
        femgpu_batch_info inf;
        inf.refinements = refinements;
        inf.inner_iterations = inner;
        inf.status = (rn <= a.tol * normF) ? 0 : 1;
        inf.pad = 0;
        inf.normF = normF;
        inf.true_rel_residual = normF > 0.0 ? rn / normF : 0.0;
        a.info[sys] = inf;
    }
}

}  // namespace

// Workgroup size from the systems' mean size when the caller leaves it to the library: 4 waves up
// to 2048 rows, 8 up to 8192, 16 above.
int wg_threads_for(long long rows, int nsys)
{
    const long long mean = rows / std::max(1, nsys);
    return mean <= 2048 ? 256 : mean <= 8192 ? 512 : 1024;
}

}  // namespace fg

using namespace fg;

extern "C" int femgpu_gmres_mixed_wg(femgpu_ctx *ctx, const femgpu_system *S, int nsys, const int *row_off,
                                     const femgpu_mixed_opts *opt, double tol, int threads, const double *d_F,
                                     double *d_X, femgpu_batch_info *info)
{
    FG_CHECK(ctx && S && S->csr && S->ilu == S->csr && !S->ebe && opt && d_F && d_X && nsys >= 1 && row_off && info,
             FEMGPU_ERR_ARG, "femgpu_gmres_mixed_wg: needs a CSR operator that is also its ILU(0) matrix");
    femgpu_csr *A = S->csr;
    const int n = A->n;
    FG_CHECK(row_off[0] == 0 && row_off[nsys] == n, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_wg: row offsets must run from 0 to n = %d", n);
    for (int s = 0; s < nsys; ++s)
        FG_CHECK(row_off[s + 1] >= row_off[s], FEMGPU_ERR_ARG, "femgpu_gmres_mixed_wg: offsets not monotone");
    FG_CHECK(opt->factor_sweeps >= 1 && opt->factor_sweeps <= 16 && opt->sweep_mode >= 1 && opt->sweep_mode <= 16,
             FEMGPU_ERR_ARG, "femgpu_gmres_mixed_wg: needs the parallel factorisation (factor_sweeps 1..16) "
             "applied by Jacobi iterations (sweep_mode 1..16)");
    FG_CHECK(opt->inner_kmax >= 2 && opt->inner_kmax <= kWgMaxBasis && opt->max_refinements >= 0, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_wg: inner basis %d outside [2, %d]", opt->inner_kmax, kWgMaxBasis);
    FG_CHECK(threads == 0 || threads == 256 || threads == 512 || threads == 1024, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_wg: threads must be 0 (automatic), 256, 512 or 1024");
    FG_CHECK(A->sell_ok && A->par_ok, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_wg: the matrix has no sliced-ELL factor layout (rows wider than 16 entries or "
             "created without the mixed-precision plans)");
    if (threads == 0) threads = wg_threads_for(n, nsys);
    const int kin = opt->inner_kmax;
    FG_TRY(launch_parilu32(ctx, A, opt->factor_sweeps));
    A->f32_factored = true;
    const size_t ldv = (static_cast<size_t>(n) + 63) / 64 * 64;
    FG_TRY(ctx->gm_basis32.alloc(ldv * (kin + 1)));
    FG_TRY(ctx->gm_w32.alloc(ldv * 3));
    FG_TRY(ctx->gm_r.alloc(static_cast<size_t>(n) + 64));
    FG_TRY(ctx->wg_info.alloc(static_cast<size_t>(nsys)));
    FG_TRY(ctx->bt_rowoff.upload(row_off, static_cast<size_t>(nsys) + 1, ctx->stream));
    WgArgs a;
    a.row_off = ctx->bt_rowoff.p;
    a.IA = A->IA.p;
    a.JA = A->JA.p;
    a.AA = A->AA.p;
    a.sptr = A->sell_ptr.p;
    a.scol = A->sell_col.p;
    a.sval = A->sell_val.p;
    a.rinfo = A->par_rinfo.p;
    a.LU = A->par_a.p;
    a.Dinv = A->Dinv32.p;
    a.F = d_F;
    a.X = d_X;
    a.R = ctx->gm_r.p;
    a.V = ctx->gm_basis32.p;
    a.T = ctx->gm_w32.p;
    a.ldv = ldv;
    a.kin = kin;
    a.jac = opt->sweep_mode;
    a.max_ref = opt->max_refinements;
    a.tol = tol;
    a.inner_tol = opt->inner_tol;
    a.info = ctx->wg_info.p;
    const int ld = kin + 1;
    const size_t lds = sizeof(double) * (static_cast<size_t>(ld) * ld + 4 * ld + 16);
    // row-width bound of the unrolled entry loops: the sliced-ELL layout holds rows of <= 16 entries
    const bool narrow = A->par_maxw <= 8;
    switch (threads) {
    case 256:
        if (narrow) k_gmres_wg<256, 8><<<nsys, 256, lds, ctx->stream>>>(a);
        else k_gmres_wg<256, 16><<<nsys, 256, lds, ctx->stream>>>(a);
        break;
    case 512:
        if (narrow) k_gmres_wg<512, 8><<<nsys, 512, lds, ctx->stream>>>(a);
        else k_gmres_wg<512, 16><<<nsys, 512, lds, ctx->stream>>>(a);
        break;
    default:
        if (narrow) k_gmres_wg<1024, 8><<<nsys, 1024, lds, ctx->stream>>>(a);
        else k_gmres_wg<1024, 16><<<nsys, 1024, lds, ctx->stream>>>(a);
        break;
    }
    FG_HIP(hipGetLastError());
    FG_HIP(hipMemcpyAsync(info, ctx->wg_info.p, sizeof(femgpu_batch_info) * nsys, hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FEMGPU_OK;
}
