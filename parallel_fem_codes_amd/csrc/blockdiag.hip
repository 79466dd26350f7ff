// Nodal block-diagonal preconditioner and block scaling for element-by-element systems with ndof
// unknowns per node (1 transport, 3 Navier-Stokes, 4 Euler):
//   * setup: the ndof x ndof diagonal block of every node, summed over the node's elements in
//     element order from the element matrices (EBE/BlockDiag_precond_EBE_setup.c:66-77,
//     EBE/BlockDiagDOF3_precond_EBE_setup.c:58-69), then inverted;
//   * apply: z = D^-1 p node by node, prescribed dofs read as 0 and not written
//     (BlockDiag_precond.c:3-29, BlockDiagDOF3_precond.c:3-27);
//   * block scaling: every element row block a replaced by D_{v_a}^-1 times the row block after
//     the boundary-condition adjustment (rows and columns of prescribed dofs zeroed, unit
//     diagonal) (Block_scaling.c:10-118).
// One thread per node (setup, apply) or per element row block (scaling); the gather lists are
// built once, so the block sums need no atomics and keep the reference's order.
#include <vector>

#include "common.h"
#include "device_util.h"

struct femgpu_blockdiag {
    femgpu_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, ndof = 0, NEQ = 0;
    fg::DevBuf<int> id;                  // nnodes*ndof: equation of the dof, NEQ when prescribed
    fg::DevBuf<int> ien;                 // nel*3
    fg::DevBuf<int> inc_ptr, inc_src, inc_dst;   // node incidences e*3+a in element order
    fg::DevBuf<double> inv;              // nnodes*ndof^2 inverse blocks (row-major)
    fg::DevBuf<float> inv32;             // their FP32 copy (mixed-precision inner solver)
    bool ready = false;
};

namespace fg {

// Inverse of a nodal block by its cofactor expansion, with the reference's own expressions
// (matrix_Inverse, EBE/BlockDiag_precond_EBE_setup.c:92-147 for 4x4; DOF3_matrix_Inverse,
// EBE/BlockDiagDOF3_precond_EBE_setup.c for 3x3), evaluated left to right with contraction off, so
// the inverse blocks are the reference's bit for bit.
template <int ND>
__device__ __forceinline__ void block_inverse(const double (&m)[ND][ND], double (&r)[ND][ND]);

template <>
__device__ __forceinline__ void block_inverse<1>(const double (&m)[1][1], double (&r)[1][1])
{
    r[0][0] = 1.0 / m[0][0];
}

template <>
__device__ __forceinline__ void block_inverse<3>(const double (&m)[3][3], double (&r)[3][3])
{
#pragma clang fp contract(off)
    const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
                 h = m[2][1], i = m[2][2];
    const double det = a * e * i - a * f * h - b * d * i + b * f * g + c * d * h - c * e * g;
    const double q = 1.0 / det;
    r[0][0] = q * (e * i - f * h);
    r[0][1] = q * (c * h - b * i);
    r[0][2] = q * (b * f - c * e);
    r[1][0] = q * (f * g - d * i);
    r[1][1] = q * (a * i - c * g);
    r[1][2] = q * (c * d - a * f);
    r[2][0] = q * (d * h - e * g);
    r[2][1] = q * (b * g - a * h);
    r[2][2] = q * (a * e - b * d);
}

template <>
__device__ __forceinline__ void block_inverse<4>(const double (&M)[4][4], double (&r)[4][4])
{
#pragma clang fp contract(off)
    const double a = M[0][0], b = M[0][1], c = M[0][2], d = M[0][3], e = M[1][0], f = M[1][1], g = M[1][2],
                 h = M[1][3], i = M[2][0], j = M[2][1], k = M[2][2], l = M[2][3], m = M[3][0], n = M[3][1],
                 o = M[3][2], p = M[3][3];
    const double det = (a * f * k * p - a * f * l * o - a * g * j * p + a * g * l * n + a * h * j * o - a * h * k * n
                        - b * e * k * p + b * e * l * o + b * g * i * p - b * g * l * m - b * h * i * o + b * h * k * m +
                        c * (e * j * p - e * l * n - f * i * p + f * l * m + h * i * n - h * j * m) +
                        d * (-e * j * o + e * k * n + f * i * o - f * k * m - g * i * n + g * j * m));
    const double q = 1.0 / det;
    r[0][0] = q * (-h * k * n + g * l * n + h * j * o - f * l * o - g * j * p + f * k * p);
    r[0][1] = q * (d * k * n - c * l * n - d * j * o + b * l * o + c * j * p - b * k * p);
    r[0][2] = q * (-d * g * n + c * h * n + d * f * o - b * h * o - c * f * p + b * g * p);
    r[0][3] = q * (d * g * j - c * h * j - d * f * k + b * h * k + c * f * l - b * g * l);
    r[1][0] = q * (h * k * m - g * l * m - h * i * o + e * l * o + g * i * p - e * k * p);
    r[1][1] = q * (-d * k * m + c * l * m + d * i * o - a * l * o - c * i * p + a * k * p);
    r[1][2] = q * (d * g * m - c * h * m - d * e * o + a * h * o + c * e * p - a * g * p);
    r[1][3] = q * (-d * g * i + c * h * i + d * e * k - a * h * k - c * e * l + a * g * l);
    r[2][0] = q * (-h * j * m + f * l * m + h * i * n - e * l * n - f * i * p + e * j * p);
    r[2][1] = q * (d * j * m - b * l * m - d * i * n + a * l * n + b * i * p - a * j * p);
    r[2][2] = q * (-d * f * m + b * h * m + d * e * n - a * h * n - b * e * p + a * f * p);
    r[2][3] = q * (d * f * i - b * h * i - d * e * j + a * h * j + b * e * l - a * f * l);
    r[3][0] = q * (g * j * m - f * k * m - g * i * n + e * k * n + f * i * o - e * j * o);
    r[3][1] = q * (-c * j * m + b * k * m + c * i * n - a * k * n - b * i * o + a * j * o);
    r[3][2] = q * (c * j * m - b * g * m - c * e * n + a * g * n + b * e * o - a * f * o);
    r[3][3] = q * (-c * f * i + b * g * i + c * e * j - a * g * j - b * e * k + a * f * k);
}

template <int ND>
__global__ __launch_bounds__(256) void k_bd_setup(int nnodes, const int *__restrict__ dst, const int *__restrict__ ptr,
                                                  const int *__restrict__ src, const double *__restrict__ Ae,
                                                  double *__restrict__ inv, float *__restrict__ inv32)
{
#pragma clang fp contract(off)
    constexpr int S = 3 * ND;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnodes) return;
    double b[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) b[i][j] = 0.0;
    for (int q = ptr[k]; q < ptr[k + 1]; ++q) {
        const int ea = src[q], e = ea / 3, a = ea % 3;
        const double *blk = Ae + static_cast<size_t>(e) * S * S + static_cast<size_t>(ND * a) * S + ND * a;
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) b[i][j] += blk[i * S + j];
    }
    double r[ND][ND];
    block_inverse<ND>(b, r);
    const size_t o = static_cast<size_t>(dst[k]) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            inv[o + i * ND + j] = r[i][j];
            inv32[o + i * ND + j] = static_cast<float>(r[i][j]);
        }
}

template <typename T, int ND>
__global__ __launch_bounds__(256) void k_bd_apply(int nnodes, int NEQ, const int *__restrict__ id,
                                                  const T *__restrict__ inv, const T *p, T *z,
                                                  const int *__restrict__ skip)
{
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int v = xcd_block() * 256 + threadIdx.x;
    if (v >= nnodes) return;
    int eq[ND];
    T x[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        eq[j] = id[v * ND + j];
        x[j] = eq[j] < NEQ ? p[eq[j]] : T(0);
    }
    const T *m = inv + static_cast<size_t>(v) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        T s = m[i * ND] * x[0];
#pragma unroll
        for (int j = 1; j < ND; ++j) s = s + m[i * ND + j] * x[j];
        if (eq[i] < NEQ) z[eq[i]] = s;
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_block_scaling(int nel, int NEQ, const int *__restrict__ ien,
                                                       const int *__restrict__ id, const double *__restrict__ inv,
                                                       double *__restrict__ Ae)
{
#pragma clang fp contract(off)
    constexpr int S = 3 * ND;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 3 * nel) return;
    const int e = t / 3, a = t % 3;
    int vtx[3];
    bool presc[S];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        vtx[b] = ien[3 * e + b];
#pragma unroll
        for (int j = 0; j < ND; ++j) presc[ND * b + j] = id[vtx[b] * ND + j] >= NEQ;
    }
    double *row = Ae + static_cast<size_t>(e) * S * S + static_cast<size_t>(ND * a) * S;
    double M[ND][S];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int c = 0; c < S; ++c) M[i][c] = presc[c] ? 0.0 : row[i * S + c];   // columns of prescribed dofs
#pragma unroll
    for (int i = 0; i < ND; ++i)
        if (presc[ND * a + i]) {   // row of a prescribed dof of this vertex: unit row
#pragma unroll
            for (int c = 0; c < S; ++c) M[i][c] = (c == ND * a + i) ? 1.0 : 0.0;
        }
    const double *m = inv + static_cast<size_t>(vtx[a]) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int c = 0; c < S; ++c) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < ND; ++j) s += m[i * ND + j] * M[j][c];
            row[i * S + c] = s;
        }
}

int blockdiag_size_check(const femgpu_blockdiag *B, int NEQ, int nnodes, int ndof)
{
    FG_CHECK(B->NEQ == NEQ && B->nnodes == nnodes && B->ndof == ndof, FEMGPU_ERR_ARG,
             "block-diagonal preconditioner: built for %d nodes x %d dofs, NEQ %d; the system has %d x %d, NEQ %d",
             B->nnodes, B->ndof, B->NEQ, nnodes, ndof, NEQ);
    return FEMGPU_OK;
}

int blockdiag_check(const femgpu_blockdiag *B, int n)
{
    FG_CHECK(B->ready, FEMGPU_ERR_STATE, "block-diagonal preconditioner: not set up");
    FG_CHECK(B->NEQ == n, FEMGPU_ERR_ARG, "block-diagonal preconditioner: NEQ %d, system size %d", B->NEQ, n);
    return FEMGPU_OK;
}

template <typename T>
static int bd_apply(femgpu_ctx *ctx, const femgpu_blockdiag *B, const T *inv, const T *p, T *z, const int *skip)
{
    FG_CHECK(B->ready, FEMGPU_ERR_STATE, "block-diagonal preconditioner: not set up");
    const int grid = xcd_grid((B->nnodes + 255) / 256);
    if (B->nnodes == 0) return FEMGPU_OK;
    switch (B->ndof) {
    case 1: k_bd_apply<T, 1><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, inv, p, z, skip); break;
    case 3: k_bd_apply<T, 3><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, inv, p, z, skip); break;
    default: k_bd_apply<T, 4><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, inv, p, z, skip); break;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

int launch_blockdiag_apply(femgpu_ctx *ctx, const femgpu_blockdiag *B, const double *p, double *z, const int *skip)
{
    return bd_apply<double>(ctx, B, B->inv.p, p, z, skip);
}

int launch_blockdiag_apply32(femgpu_ctx *ctx, const femgpu_blockdiag *B, const float *p, float *z, const int *skip)
{
    return bd_apply<float>(ctx, B, B->inv32.p, p, z, skip);
}

}  // namespace fg

using namespace fg;

extern "C" int femgpu_blockdiag_create(femgpu_ctx *ctx, int nnodes, int nel, int ndof, int NEQ, const int *ien,
                                       const int *type, const int *invP, femgpu_blockdiag **out)
{
    FG_CHECK(ctx && out && nnodes >= 0 && nel >= 0 && NEQ >= 0 && ien && type && invP &&
                 (ndof == 1 || ndof == 3 || ndof == 4),
             FEMGPU_ERR_ARG, "femgpu_blockdiag_create: bad arguments");
    std::vector<int> id(static_cast<size_t>(nnodes) * ndof);
    for (size_t k = 0; k < id.size(); ++k) {
        FG_CHECK(type[k] == 0 || (invP[k] >= 0 && invP[k] < NEQ), FEMGPU_ERR_ARG,
                 "femgpu_blockdiag_create: dof %zu has invP_id %d outside [0, NEQ)", k, invP[k]);
        id[k] = type[k] == 0 ? NEQ : invP[k];
    }
    std::vector<std::pair<int, int>> pairs;
    pairs.reserve(static_cast<size_t>(nel) * 3);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const int v = ien[3 * static_cast<size_t>(e) + a];
            FG_CHECK(v >= 0 && v < nnodes, FEMGPU_ERR_ARG, "femgpu_blockdiag_create: vertex %d out of range", v);
            pairs.emplace_back(v, 3 * e + a);
        }
    std::vector<int> ptr, src, dst;
    csr_map(nnodes, pairs, ptr, src, dst);
    auto *B = new femgpu_blockdiag();
    B->ctx = ctx;
    B->nnodes = nnodes;
    B->nel = nel;
    B->ndof = ndof;
    B->NEQ = NEQ;
    hipStream_t s = ctx->stream;
    int rc = B->id.upload(id.data(), id.size(), s);
    if (!rc) rc = B->ien.upload(ien, static_cast<size_t>(nel) * 3, s);
    if (!rc) rc = B->inc_ptr.upload(ptr.data(), ptr.size(), s);
    if (!rc) rc = B->inc_src.upload(src.data(), src.size(), s);
    if (!rc) rc = B->inc_dst.upload(dst.data(), dst.size(), s);
    if (!rc) rc = B->inv.alloc(static_cast<size_t>(nnodes) * ndof * ndof);
    if (!rc) rc = B->inv32.alloc(static_cast<size_t>(nnodes) * ndof * ndof);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = FEMGPU_ERR_HIP;
    if (rc) {
        delete B;
        return rc;
    }
    *out = B;
    return FEMGPU_OK;
}

extern "C" int femgpu_blockdiag_destroy(femgpu_blockdiag *B)
{
    delete B;
    return FEMGPU_OK;
}

extern "C" int femgpu_blockdiag_setup(femgpu_blockdiag *B, const double *d_Ae)
{
    FG_CHECK(B && d_Ae, FEMGPU_ERR_ARG, "femgpu_blockdiag_setup: bad arguments");
    femgpu_ctx *ctx = B->ctx;
    const int grid = (B->nnodes + 255) / 256;
    if (B->nnodes > 0) {
        switch (B->ndof) {
        case 1: k_bd_setup<1><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p, B->inv32.p); break;
        case 3: k_bd_setup<3><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p, B->inv32.p); break;
        default: k_bd_setup<4><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p, B->inv32.p); break;
        }
        FG_HIP(hipGetLastError());
    }
    B->ready = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_blockdiag_apply(femgpu_blockdiag *B, const double *d_p, double *d_z)
{
    FG_CHECK(B && d_p && d_z, FEMGPU_ERR_ARG, "femgpu_blockdiag_apply: bad arguments");
    return launch_blockdiag_apply(B->ctx, B, d_p, d_z, nullptr);
}

extern "C" int femgpu_block_scaling(femgpu_blockdiag *B, double *d_Ae, double *d_F)
{
    FG_CHECK(B && d_Ae, FEMGPU_ERR_ARG, "femgpu_block_scaling: bad arguments");
    FG_TRY(femgpu_blockdiag_setup(B, d_Ae));   // Block_scaling.c:20, from the unscaled matrices
    femgpu_ctx *ctx = B->ctx;
    if (d_F) FG_TRY(launch_blockdiag_apply(ctx, B, d_F, d_F, nullptr));   // F := D^-1 F (the setup's :84)
    const int grid = (3 * B->nel + 255) / 256;
    if (B->nel > 0) {
        switch (B->ndof) {
        case 1: k_block_scaling<1><<<grid, 256, 0, ctx->stream>>>(B->nel, B->NEQ, B->ien.p, B->id.p, B->inv.p, d_Ae); break;
        case 3: k_block_scaling<3><<<grid, 256, 0, ctx->stream>>>(B->nel, B->NEQ, B->ien.p, B->id.p, B->inv.p, d_Ae); break;
        default: k_block_scaling<4><<<grid, 256, 0, ctx->stream>>>(B->nel, B->NEQ, B->ien.p, B->id.p, B->inv.p, d_Ae); break;
        }
        FG_HIP(hipGetLastError());
    }
    return FEMGPU_OK;
}
