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

struct femgpu_blockdiag {
    femgpu_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, ndof = 0, NEQ = 0;
    fg::DevBuf<int> id;                  // nnodes*ndof: equation of the dof, NEQ when prescribed
    fg::DevBuf<int> ien;                 // nel*3
    fg::DevBuf<int> inc_ptr, inc_src, inc_dst;   // node incidences e*3+a in element order
    fg::DevBuf<double> inv;              // nnodes*ndof^2 inverse blocks (row-major)
    bool ready = false;
};

namespace fg {

// Inverse of a small dense block by Gauss-Jordan elimination with partial pivoting (the blocks
// are the diagonal blocks of stabilised element matrices; pivoting keeps the 4x4 Euler blocks
// accurate when a leading entry is small).
template <int ND>
__device__ __forceinline__ void block_inverse(double (&a)[ND][ND], double (&inv)[ND][ND])
{
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) inv[i][j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int c = 0; c < ND; ++c) {
        int piv = c;
        double best = fabs(a[c][c]);
#pragma unroll
        for (int r = c + 1; r < ND; ++r)
            if (fabs(a[r][c]) > best) {
                best = fabs(a[r][c]);
                piv = r;
            }
#pragma unroll
        for (int r = c + 1; r < ND; ++r)
            if (r == piv)
#pragma unroll
                for (int j = 0; j < ND; ++j) {
                    double t = a[c][j];
                    a[c][j] = a[r][j];
                    a[r][j] = t;
                    t = inv[c][j];
                    inv[c][j] = inv[r][j];
                    inv[r][j] = t;
                }
        const double d = 1.0 / a[c][c];
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            a[c][j] *= d;
            inv[c][j] *= d;
        }
#pragma unroll
        for (int r = 0; r < ND; ++r)
            if (r != c) {
                const double f = a[r][c];
#pragma unroll
                for (int j = 0; j < ND; ++j) {
                    a[r][j] -= f * a[c][j];
                    inv[r][j] -= f * inv[c][j];
                }
            }
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_bd_setup(int nnodes, const int *__restrict__ dst, const int *__restrict__ ptr,
                                                  const int *__restrict__ src, const double *__restrict__ Ae,
                                                  double *__restrict__ inv)
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
    double *o = inv + static_cast<size_t>(dst[k]) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) o[i * ND + j] = r[i][j];
}

template <int ND>
__global__ __launch_bounds__(256) void k_bd_apply(int nnodes, int NEQ, const int *__restrict__ id,
                                                  const double *__restrict__ inv, const double *p, double *z,
                                                  const int *__restrict__ skip)
{
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nnodes) return;
    int eq[ND];
    double x[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        eq[j] = id[v * ND + j];
        x[j] = eq[j] < NEQ ? p[eq[j]] : 0.0;
    }
    const double *m = inv + static_cast<size_t>(v) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double s = m[i * ND] * x[0];
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

int blockdiag_check(const femgpu_blockdiag *B, int n)
{
    FG_CHECK(B->ready, FEMGPU_ERR_STATE, "block-diagonal preconditioner: not set up");
    FG_CHECK(B->NEQ == n, FEMGPU_ERR_ARG, "block-diagonal preconditioner: NEQ %d, system size %d", B->NEQ, n);
    return FEMGPU_OK;
}

int launch_blockdiag_apply(femgpu_ctx *ctx, const femgpu_blockdiag *B, const double *p, double *z, const int *skip)
{
    FG_CHECK(B->ready, FEMGPU_ERR_STATE, "block-diagonal preconditioner: not set up");
    const int grid = (B->nnodes + 255) / 256;
    if (B->nnodes == 0) return FEMGPU_OK;
    switch (B->ndof) {
    case 1: k_bd_apply<1><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, B->inv.p, p, z, skip); break;
    case 3: k_bd_apply<3><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, B->inv.p, p, z, skip); break;
    default: k_bd_apply<4><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->NEQ, B->id.p, B->inv.p, p, z, skip); break;
    }
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
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
        case 1: k_bd_setup<1><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p); break;
        case 3: k_bd_setup<3><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p); break;
        default: k_bd_setup<4><<<grid, 256, 0, ctx->stream>>>(B->nnodes, B->inc_dst.p, B->inc_ptr.p, B->inc_src.p, d_Ae, B->inv.p); break;
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
