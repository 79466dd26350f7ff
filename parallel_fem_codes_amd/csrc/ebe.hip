// Element-by-element operator (ebemv.c:4-64, ebemvNDOF3.c:3-92, ebemvNDOF4.c:4-121): Q = sum_e A_e P|_e
// for s x s element blocks (s = 3 ndof, ndof = 1 transport, 3 Navier-Stokes, 4 Euler).
//
// The product is gathered per node instead of scattered per element. Element row ndof*a + d of
// element e belongs to dof d of its vertex a; a node's rows over all its elements form one group,
// summed by one thread in element order:
//   Q[eq_d] = sum over the node's elements e (ascending) of  sum_b A_e[ndof*a + d][b] * P[lm_e[b]]
// with the row sum left to right and P at the Dirichlet sink read as 0 -- the reference's own
// operations in its own order (its element loop adds the row sums of element e to Q in ascending
// e), so Q is bit-identical to its element loop, and written exactly once: no colouring, atomics or
// read-modify-write. Groups go in slices of 64, one lane each; a slice stores its incidences
// column-major (value (d, b) of incidence row r, lane l at ((r*ndof + d)*s + b)*64 + l), so every
// load of a wavefront is one contiguous run; the element rows of prescribed dofs are not stored.
// Columns are kept as int16 offsets from the node's first unknown when the mesh numbering allows
// (2 bytes per entry), else as int32 unknowns.
#include <climits>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

constexpr short kSink16 = -32768;   // int16 column of the Dirichlet sink

template <typename CT>
__device__ __forceinline__ int ebe_col(CT c, int key, int NEQ);
template <>
__device__ __forceinline__ int ebe_col<short>(short c, int key, int)
{
    return c == kSink16 ? -1 : key + c;
}
template <>
__device__ __forceinline__ int ebe_col<int>(int c, int, int NEQ)
{
    return c < NEQ ? c : -1;
}

// One thread per node group. The loads of U incidences are issued together (incidences past the
// list re-read the first of the batch and are not added), then summed in list order. No
// contraction into fused multiply-adds: products and sums round as the reference's.
template <typename T, typename CT, int ND, int U>
__global__ __launch_bounds__(256) void k_ebe_gather(int ngroups, int NEQ, const int *__restrict__ iptr,
                                                    const int *__restrict__ len, const int *__restrict__ eq,
                                                    const int *__restrict__ key, const CT *__restrict__ col,
                                                    const T *__restrict__ val, const T *__restrict__ P,
                                                    T *__restrict__ Q, const int *__restrict__ skip)
{
#pragma clang fp contract(off)
    constexpr int S = 3 * ND;
    if (skip && *skip) return;
    const int g = xcd_block() * 256 + static_cast<int>(threadIdx.x);
    if (g >= ngroups) return;
    const int sl = g >> 6, l = g & 63;
    const int r0 = iptr[sl], n = len[g];
    const int kg = key ? key[g] : 0;
    T acc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[d] = T(0);
    for (int j = 0; j < n; j += U) {
        T a[U][ND][S], p[U][S];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t r = static_cast<size_t>(r0 + (j + u < n ? j + u : j));
#pragma unroll
            for (int b = 0; b < S; ++b) {
                const int idx = ebe_col<CT>(col[(r * S + b) * 64 + l], kg, NEQ);
                const T pv = P[idx >= 0 ? idx : 0];
                p[u][b] = idx >= 0 ? pv : T(0);
            }
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int b = 0; b < S; ++b) a[u][d][b] = val[((r * ND + d) * S + b) * 64 + l];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j + u < n) {
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    T t = a[u][d][0] * p[u][0];
#pragma unroll
                    for (int b = 1; b < S; ++b) t = t + a[u][d][b] * p[u][b];
                    acc[d] = acc[d] + t;
                }
            }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const int q = eq[(static_cast<size_t>(sl) * ND + d) * 64 + l];
        if (q < NEQ) Q[q] = acc[d];
    }
}

template <typename T>
__global__ void k_ebe_zero(int n, T *__restrict__ q, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) q[i] = T(0);
}

// Element matrices (element order, ebe_assembly.c layout) into the gathered layout, FP64 and FP32.
template <int ND>
__global__ __launch_bounds__(256) void k_ebe_fill(int nslots, const int *__restrict__ src, const double *__restrict__ Ae,
                                                  double *__restrict__ val, float *__restrict__ val32)
{
    constexpr int S = 3 * ND;
    const int q = blockIdx.x * 256 + static_cast<int>(threadIdx.x);
    if (q >= nslots) return;
    const size_t r = static_cast<size_t>(q >> 6);
    const int l = q & 63, base = src[q];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int b = 0; b < S; ++b) {
            const double v = base >= 0 ? Ae[static_cast<size_t>(base) + d * S + b] : 0.0;
            const size_t o = ((r * ND + d) * S + b) * 64 + l;
            val[o] = v;
            val32[o] = static_cast<float>(v);
        }
}

template <typename T, typename CT, int ND, int U>
static void gather_launch(femgpu_ctx *ctx, const femgpu_ebe *E, const T *vals, const T *P, T *Q, const int *skip)
{
    const int grid = xcd_grid((E->ngroups + 255) / 256);
    const CT *col;
    if constexpr (sizeof(CT) == 2) col = E->col16v.p;
    else col = E->col32v.p;
    k_ebe_gather<T, CT, ND, U><<<grid, 256, 0, ctx->stream>>>(E->ngroups, E->NEQ, E->iptr.p, E->len.p, E->eq.p,
                                                              E->col16 ? E->key.p : nullptr, col, vals, P, Q, skip);
}

template <typename T>
static int ebe_product(femgpu_ctx *ctx, const femgpu_ebe *E, const T *vals, const T *P, T *Q, const int *skip)
{
    FG_CHECK(E->has_values, FEMGPU_ERR_STATE, "femgpu_ebemv: element matrices not set");
    if (E->zero_q && E->NEQ > 0) {
        k_ebe_zero<T><<<std::min(4096, (E->NEQ + 255) / 256), 256, 0, ctx->stream>>>(E->NEQ, Q, skip);
        FG_HIP(hipGetLastError());
    }
    if (E->ngroups == 0) return FEMGPU_OK;
    const bool c16 = E->col16 != 0;
    switch (E->ndof) {
    case 1:
        if (c16) gather_launch<T, short, 1, 6>(ctx, E, vals, P, Q, skip);
        else gather_launch<T, int, 1, 6>(ctx, E, vals, P, Q, skip);
        break;
    case 3:
        if (c16) gather_launch<T, short, 3, 1>(ctx, E, vals, P, Q, skip);
        else gather_launch<T, int, 3, 1>(ctx, E, vals, P, Q, skip);
        break;
    default:
        if (c16) gather_launch<T, short, 4, 1>(ctx, E, vals, P, Q, skip);
        else gather_launch<T, int, 4, 1>(ctx, E, vals, P, Q, skip);
        break;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int launch_ebemv(femgpu_ctx *ctx, const femgpu_ebe *E, const double *P, double *Q, const int *skip)
{
    return ebe_product<double>(ctx, E, E->val.p, P, Q, skip);
}

int launch_ebemv32(femgpu_ctx *ctx, const femgpu_ebe *E, const float *P, float *Q, const int *skip)
{
    return ebe_product<float>(ctx, E, E->val32.p, P, Q, skip);
}

}  // namespace fg

using namespace fg;

extern "C" int femgpu_ebe_create(femgpu_ctx *ctx, int nel, int s, int NEQ, const int *lm, femgpu_ebe **out)
{
    FG_CHECK(ctx && out && nel >= 0 && NEQ >= 0 && lm && (s == 3 || s == 9 || s == 12), FEMGPU_ERR_ARG,
             "femgpu_ebe_create: bad arguments (s must be 3, 9 or 12)");
    FG_CHECK(static_cast<long long>(nel) * s * s < INT_MAX, FEMGPU_ERR_ARG, "femgpu_ebe_create: %d elements of %dx%d "
             "exceed the int32 element-matrix index", nel, s, s);
    const size_t total = static_cast<size_t>(nel) * s;
    for (size_t q = 0; q < total; ++q)
        FG_CHECK(lm[q] >= 0 && lm[q] <= NEQ, FEMGPU_ERR_ARG, "femgpu_ebe_create: lm entry %d out of [0, NEQ]", lm[q]);
    const int nd = s / 3;
    // incidence (e, a) -> its node's first unknown (the group key); nodes with every dof prescribed
    // contribute no rows
    std::vector<int> kinc(static_cast<size_t>(nel) * 3, -1), cnt(static_cast<size_t>(NEQ) + 1, 0);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const int *L = lm + static_cast<size_t>(e) * s + nd * a;
            int k = -1;
            for (int d = 0; d < nd && k < 0; ++d)
                if (L[d] < NEQ) k = L[d];
            kinc[3 * static_cast<size_t>(e) + a] = k;
            if (k >= 0) ++cnt[k];
        }
    std::vector<int> gid(static_cast<size_t>(NEQ) + 1, -1), gkey;
    for (int k = 0; k < NEQ; ++k)
        if (cnt[k] > 0) {
            gid[k] = static_cast<int>(gkey.size());
            gkey.push_back(k);
        }
    const int ng = static_cast<int>(gkey.size());
    std::vector<int> gptr(static_cast<size_t>(ng) + 1, 0);
    for (int g = 0; g < ng; ++g) gptr[g + 1] = gptr[g] + cnt[gkey[g]];
    std::vector<int> inc(gptr[ng]), fillp(gptr.begin(), gptr.end() - 1);
    for (size_t ea = 0; ea < kinc.size(); ++ea)
        if (kinc[ea] >= 0) inc[fillp[gid[kinc[ea]]]++] = static_cast<int>(ea);   // element order kept
    // the node's unknowns, identical in every element at the node
    const int nslices = (ng + 63) / 64;
    std::vector<int> eq(static_cast<size_t>(nslices) * nd * 64, NEQ);
    std::vector<char> covered(static_cast<size_t>(NEQ) + 1, 0);
    for (int g = 0; g < ng; ++g) {
        const int e0 = inc[gptr[g]] / 3, a0 = inc[gptr[g]] % 3;
        for (int q = gptr[g]; q < gptr[g + 1]; ++q) {
            const int e = inc[q] / 3, a = inc[q] % 3;
            for (int d = 0; d < nd; ++d)
                FG_CHECK(lm[static_cast<size_t>(e) * s + nd * a + d] == lm[static_cast<size_t>(e0) * s + nd * a0 + d],
                         FEMGPU_ERR_ARG, "femgpu_ebe_create: element %d and %d give one node different unknowns", e0, e);
        }
        for (int d = 0; d < nd; ++d) {
            const int u = lm[static_cast<size_t>(e0) * s + nd * a0 + d];
            eq[(static_cast<size_t>(g >> 6) * nd + d) * 64 + (g & 63)] = u;
            covered[u] = 1;
        }
    }
    bool zero_q = false;
    for (int k = 0; k < NEQ; ++k) zero_q = zero_q || !covered[k];
    std::vector<int> iptr(static_cast<size_t>(nslices) + 1, 0), len(ng);
    int maxinc = 0;
    for (int sl = 0; sl < nslices; ++sl) {
        int w = 0;
        for (int g = 64 * sl; g < std::min(ng, 64 * sl + 64); ++g) {
            len[g] = gptr[g + 1] - gptr[g];
            w = std::max(w, len[g]);
        }
        maxinc = std::max(maxinc, w);
        iptr[sl + 1] = iptr[sl] + w;
    }
    const int nrows = iptr[nslices];
    const size_t nslots = static_cast<size_t>(nrows) * 64;
    FG_CHECK(nslots * s * nd < (size_t{1} << 40) && nslots < static_cast<size_t>(INT_MAX), FEMGPU_ERR_ARG,
             "femgpu_ebe_create: operator too large");
    std::vector<int> src(nslots, -1), c32(nslots * s, NEQ);
    std::vector<short> c16(nslots * s, kSink16);
    bool fits16 = true;
    for (int g = 0; g < ng; ++g) {
        const int sl = g >> 6, l = g & 63;
        for (int j = 0; j < len[g]; ++j) {
            const int ea = inc[gptr[g] + j], e = ea / 3, a = ea % 3;
            const size_t r = static_cast<size_t>(iptr[sl] + j);
            src[r * 64 + l] = e * s * s + nd * a * s;
            for (int b = 0; b < s; ++b) {
                const int c = lm[static_cast<size_t>(e) * s + b];
                c32[(r * s + b) * 64 + l] = c;
                if (c < NEQ) {
                    const long dlt = static_cast<long>(c) - gkey[g];
                    if (dlt < -32767 || dlt > 32767) fits16 = false;
                    else c16[(r * s + b) * 64 + l] = static_cast<short>(dlt);
                }
            }
        }
    }
    auto *E = new femgpu_ebe();
    E->ctx = ctx;
    E->nel = nel;
    E->s = s;
    E->ndof = nd;
    E->NEQ = NEQ;
    E->ngroups = ng;
    E->nrows = nrows;
    E->maxinc = maxinc;
    E->col16 = fits16 ? 1 : 0;
    E->zero_q = zero_q;
    E->hiptr = iptr;
    hipStream_t st = ctx->stream;
    int rc = E->iptr.upload(iptr.data(), iptr.size(), st);
    if (!rc) rc = E->len.upload(len.data(), len.size(), st);
    if (!rc) rc = E->eq.upload(eq.data(), eq.size(), st);
    if (!rc) rc = E->src.upload(src.data(), src.size(), st);
    if (!rc && fits16) rc = E->key.upload(gkey.data(), gkey.size(), st);
    if (!rc && fits16) rc = E->col16v.upload(c16.data(), c16.size(), st);
    if (!rc && !fits16) rc = E->col32v.upload(c32.data(), c32.size(), st);
    if (!rc) rc = E->val.alloc(nslots * s * nd);
    if (!rc) rc = E->val32.alloc(nslots * s * nd);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = FEMGPU_ERR_HIP;   // host arrays end here
    if (rc) {
        delete E;
        return rc;
    }
    *out = E;
    return FEMGPU_OK;
}

extern "C" int femgpu_ebe_destroy(femgpu_ebe *E)
{
    delete E;
    return FEMGPU_OK;
}

extern "C" int femgpu_ebe_set_values(femgpu_ebe *E, const double *A, int on_device)
{
    FG_CHECK(E && A, FEMGPU_ERR_ARG, "femgpu_ebe_set_values: bad arguments");
    femgpu_ctx *ctx = E->ctx;
    const size_t total = static_cast<size_t>(E->nel) * E->s * E->s;
    DevBuf<double> tmp;
    const double *dA = A;
    if (!on_device) {
        FG_TRY(tmp.upload(A, total, ctx->stream));
        dA = tmp.p;
    }
    const int nslots = E->nrows * 64;
    if (nslots > 0) {
        const int grid = (nslots + 255) / 256;
        switch (E->ndof) {
        case 1: k_ebe_fill<1><<<grid, 256, 0, ctx->stream>>>(nslots, E->src.p, dA, E->val.p, E->val32.p); break;
        case 3: k_ebe_fill<3><<<grid, 256, 0, ctx->stream>>>(nslots, E->src.p, dA, E->val.p, E->val32.p); break;
        default: k_ebe_fill<4><<<grid, 256, 0, ctx->stream>>>(nslots, E->src.p, dA, E->val.p, E->val32.p); break;
        }
        FG_HIP(hipGetLastError());
    }
    if (!on_device) FG_HIP(hipStreamSynchronize(ctx->stream));
    E->has_values = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_ebe_get_info(femgpu_ebe *E, femgpu_ebe_info *info)
{
    FG_CHECK(E && info, FEMGPU_ERR_ARG, "femgpu_ebe_get_info: bad arguments");
    const long long slots = static_cast<long long>(E->nrows) * 64, s = E->s, nd = E->ndof;
    const long long nsl = static_cast<long long>(E->hiptr.size()) - 1;
    const long long meta = 4LL * E->ngroups * (E->col16 ? 2 : 1) + 4 * (nsl + 1) + 4 * nsl * nd * 64;
    const long long cols = slots * s * (E->col16 ? 2 : 4);
    info->groups = E->ngroups;
    info->max_incidences = E->maxinc;
    info->column_bytes = E->col16 ? 2 : 4;
    info->pad = 0;
    info->bytes_f64 = slots * nd * s * 8 + cols + meta + 16LL * E->NEQ;
    info->bytes_f32 = slots * nd * s * 4 + cols + meta + 8LL * E->NEQ;
    return FEMGPU_OK;
}

extern "C" int femgpu_ebemv(femgpu_ebe *E, const double *d_P, double *d_Q)
{
    FG_CHECK(E && d_P && d_Q && d_P != d_Q, FEMGPU_ERR_ARG, "femgpu_ebemv: bad arguments");
    return launch_ebemv(E->ctx, E, d_P, d_Q);
}

extern "C" int femgpu_ebemv_f32(femgpu_ebe *E, const float *d_P, float *d_Q)
{
    FG_CHECK(E && d_P && d_Q && d_P != d_Q, FEMGPU_ERR_ARG, "femgpu_ebemv_f32: bad arguments");
    return launch_ebemv32(E->ctx, E, d_P, d_Q);
}
