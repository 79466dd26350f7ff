// Per-element SUPG/PSPG/LSIC assembly of the steady incompressible Navier-Stokes equations
// (P1 u, v, p; 9x9 element blocks) on the device: the Picard tangent and residual of
// Build_K_F_SUPG_PSPG.c:8-671 (the branch its call counter selects below 50000, :466-566) with the
// nodal state of eval_U.c:3-38.
//
// One thread per element. Vertex a has the edge vectors b[a] = (y23, y31, y12)[a] and
// c[a] = (x32, x13, x21)[a]; each entry of the element matrix is the reference's expression for
// its vertex pair, evaluated in the reference's operation order with contraction off (for the
// symmetric terms the reference evaluates the (min, max) pair and reuses it, and so does this
// kernel), so element matrices, residuals, CSR values and right-hand sides are bit-identical.
// Global accumulation is a gather in element order, as for the transport assembly.
#include "common.h"
#include "device_util.h"

namespace fg {

struct NsIn {
    const double *x, *y, *presc, *u, *fx, *fy;
    const int *type, *invP, *ien;
    double Re;
};

__device__ __forceinline__ double ns_dof(const NsIn &in, int v, int d)
{
    const int k = 3 * v + d;
    return in.type[k] > 0 ? in.u[in.invP[k]] : in.presc[k];
}

__global__ __launch_bounds__(256) void k_elem_ns(int nel, NsIn in, double *__restrict__ Ke, double *__restrict__ Res)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nel) return;
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    int v[3];
    for (int a = 0; a < 3; ++a) v[a] = in.ien[3 * e + a];
    const double X0 = in.x[v[0]], X1 = in.x[v[1]], X2 = in.x[v[2]];
    const double Y0 = in.y[v[0]], Y1 = in.y[v[1]], Y2 = in.y[v[2]];
    const double b[3] = {Y1 - Y2, Y2 - Y0, Y0 - Y1};
    const double c[3] = {X2 - X1, X0 - X2, X1 - X0};
    const double TwoA = fabs(c[2] * b[1] - c[1] * b[2]);
    const double Area = 0.5 * TwoA, invArea = 1.0 / Area;
    double uxv[3], uyv[3], pv[3];
    for (int a = 0; a < 3; ++a) {
        uxv[a] = ns_dof(in, v[a], 0);
        uyv[a] = ns_dof(in, v[a], 1);
        pv[a] = ns_dof(in, v[a], 2);
    }
    const double ux = third * (uxv[0] + uxv[1] + uxv[2]);
    const double uy = third * (uyv[0] + uyv[1] + uyv[2]);
    const double duxdx = 0.5 * invArea * (uxv[0] * b[0] + uxv[1] * b[1] + uxv[2] * b[2]);
    const double duxdy = 0.5 * invArea * (uxv[0] * c[0] + uxv[1] * c[1] + uxv[2] * c[2]);
    const double duydx = 0.5 * invArea * (uyv[0] * b[0] + uyv[1] * b[1] + uyv[2] * b[2]);
    const double duydy = 0.5 * invArea * (uyv[0] * c[0] + uyv[1] * c[1] + uyv[2] * c[2]);
    const double duxduy = uxv[0] * b[0] + uyv[0] * c[0] + uxv[1] * b[1] + uyv[1] * c[1] + uxv[2] * b[2] +
                          uyv[2] * c[2];
    const double p = third * (pv[0] + pv[1] + pv[2]);
    const double dpdx = 0.5 * invArea * (pv[0] * b[0] + pv[1] * b[1] + pv[2] * b[2]);
    const double dpdy = 0.5 * invArea * (pv[0] * c[0] + pv[1] * c[1] + pv[2] * c[2]);
    double C[3];
    for (int a = 0; a < 3; ++a) C[a] = 0.5 * invArea * (ux * b[a] + uy * c[a]);
    // tau_SUPG = tau_PSPG and tau_LSIC (:118-131), rho = 1, viscosity 1/Re
    const double h = sqrt(TwoA), visc = 1. / in.Re, rho = 1.;
    const double invrho = 1.0 / rho, nu = visc / rho;
    const double unorm = sqrt(ux * ux + uy * uy);
    double auxtau = (2.0 * unorm / h) * (2.0 * unorm / h) + 9.0 * (4.0 * nu / (h * h)) * (4.0 * nu / (h * h));
    auxtau = sqrt(auxtau);
    const double tau = 1.0 / auxtau;
    const double tau_l = sqrt(nu * nu + (0.5 * h * unorm) * (0.5 * h * unorm));
    const double P6[6] = {b[0], c[0], b[1], c[1], b[2], c[2]};
    double *K = Ke + 81 * static_cast<size_t>(e);
    for (int a = 0; a < 3; ++a) {
        for (int bb = 0; bb < 3; ++bb) {
            const int lo = a < bb ? a : bb, hi = a < bb ? bb : a;
            const double Nb = rho * third * Area * C[bb];
            const double Nd = rho * tau * Area * C[lo] * C[hi];
            const double Kuu = visc * 0.25 * invArea * (2 * b[lo] * b[hi] + c[lo] * c[hi]);
            const double Kvv = visc * 0.25 * invArea * (b[lo] * b[hi] + 2 * c[lo] * c[hi]);
            const double Kuv = visc * 0.25 * invArea * (c[a] * b[bb]);
            const double Kvu = visc * 0.25 * invArea * (b[a] * c[bb]);
            const double ks = tau_l * rho * 0.25 * invArea;
            const double Ks_uu = ks * P6[2 * bb] * P6[2 * a], Ks_uv = ks * P6[2 * bb + 1] * P6[2 * a];
            const double Ks_vu = ks * P6[2 * bb] * P6[2 * a + 1], Ks_vv = ks * P6[2 * bb + 1] * P6[2 * a + 1];
            const double Gdu_ab = 0.5 * tau * C[a] * b[bb], Gdv_ab = 0.5 * tau * C[a] * c[bb];
            const double Gdu_ba = 0.5 * tau * C[bb] * b[a], Gdv_ba = 0.5 * tau * C[bb] * c[a];
            const double Gphi = invrho * 0.25 * tau * invArea * (b[lo] * b[hi] + c[lo] * c[hi]);
            double *r0 = K + 27 * a + 3 * bb;
            r0[0] = Nb + Nd + Kuu + Ks_uu;
            r0[1] = Kuv + Ks_uv;
            r0[2] = -(sixth * b[a] + Gdu_ab);
            r0[9] = Kvu + Ks_vu;
            r0[10] = Nb + Nd + Kvv + Ks_vv;
            r0[11] = -(sixth * c[a] + Gdv_ab);
            r0[18] = sixth * b[bb] + Gdu_ba;
            r0[19] = sixth * c[bb] + Gdv_ba;
            r0[20] = Gphi;
        }
    }
    // loads and residual (:392-464)
    const double fxB = in.fx ? third * (in.fx[v[0]] + in.fx[v[1]] + in.fx[v[2]]) : 0.0;
    const double fyB = in.fy ? third * (in.fy[v[0]] + in.fy[v[1]] + in.fy[v[2]]) : 0.0;
    const double f1 = rho * third * Area * fxB, f2 = rho * third * Area * fyB;
    const double cvx = duxdx * ux + duxdy * uy, cvy = duydx * ux + duydy * uy;
    const double Nvx = rho * third * Area * cvx, Nvy = rho * third * Area * cvy;
    const double GTv = third * Area * (duxdx + duydy);
    double *R = Res + 9 * static_cast<size_t>(e);
    for (int a = 0; a < 3; ++a) {
        const double fdx = rho * tau * Area * C[a] * fxB, fdy = rho * tau * Area * C[a] * fyB;
        const double fphi = tau * Area * (b[a] * fxB + c[a] * fyB);
        const double Ndx = rho * tau * Area * C[a] * cvx, Ndy = rho * tau * Area * C[a] * cvy;
        const double Kvx = 0.5 * visc * (2 * b[a] * duxdx + c[a] * (duxdy + duydx));
        const double Kvy = 0.5 * visc * (2 * c[a] * duydy + b[a] * (duxdy + duydx));
        const double Ksx = tau_l * rho * 0.25 * invArea * b[a] * duxduy;
        const double Ksy = tau_l * rho * 0.25 * invArea * c[a] * duxduy;
        const double Gvx = 0.5 * p * b[a], Gvy = 0.5 * p * c[a];
        const double Gdx = tau * Area * C[a] * dpdx, Gdy = tau * Area * C[a] * dpdy;
        const double Nphi = 0.5 * tau * ((duxdx * ux + duxdy * uy) * b[a] + (duydx * ux + duydy * uy) * c[a]);
        const double Gphiv = 0.5 * tau * invrho * (b[a] * dpdx + c[a] * dpdy);
        R[3 * a] = (f1 + fdx) - (Nvx + Ndx + Kvx + Ksx - (Gvx + Gdx));
        R[3 * a + 1] = (f2 + fdy) - (Nvy + Ndy + Kvy + Ksy - (Gvy + Gdy));
        R[3 * a + 2] = fphi - (GTv + Nphi + Gphiv);
    }
}

// F[I] = sum of the residual entries of unknown I in element order (F[lm] += Res, :655-658);
// F[NEQ] = 0. src entries index the per-element residual array (e*s+i).
__global__ void k_rhs_gather(int NEQ, const int *__restrict__ dst, const int *__restrict__ ptr,
                             const int *__restrict__ src, const double *__restrict__ Res, double *__restrict__ F)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > NEQ) return;
    if (k == NEQ) {
        F[NEQ] = 0.0;
        return;
    }
    F[dst[k]] = gather_sum<8>(src, Res, ptr[k], ptr[k + 1]);
}

__global__ void k_aa_gather_blocks(int nnz, const int *__restrict__ dst, const int *__restrict__ ptr,
                                   const int *__restrict__ src, const double *__restrict__ Ke, double *__restrict__ AA)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    if (k == nnz) {
        AA[nnz] = 0.0;
        return;
    }
    AA[dst[k]] = gather_sum<8>(src, Ke, ptr[k], ptr[k + 1]);
}

}  // namespace fg

using namespace fg;

namespace fg {

// Device mesh for element blocks of s = 3*ndof dofs (ndof = 3: Navier-Stokes, 4: Euler): per-dof
// Type/invP_id/prescribed values, lm[nel*s], the residual gather map (unknown -> e*s+i in element
// order) and, with a slot table scheme[nel*s*s], the CSR value gather map (slot -> e*s*s+q).
int block_mesh_create(femgpu_ctx *ctx, int ndof, int nnodes, int nel, int neq, const double *x, const double *y,
                      const int *type, const int *invP, const double *presc, const int *ien, const int *lm,
                      const int *scheme, int nnz, femgpu_mesh **out)
{
    const int sb = 3 * ndof, ss = sb * sb;
    FG_CHECK(ctx && nnodes > 0 && nel >= 0 && neq >= 0 && x && y && type && invP && presc && ien && lm && out,
             FEMGPU_ERR_ARG, "mesh create: bad arguments");
    for (size_t q = 0; q < static_cast<size_t>(nel) * 3; ++q)
        FG_CHECK(ien[q] >= 0 && ien[q] < nnodes, FEMGPU_ERR_ARG, "mesh create: vertex %d out of range", ien[q]);
    for (size_t k = 0; k < static_cast<size_t>(nnodes) * ndof; ++k)
        FG_CHECK(type[k] == 0 || (invP[k] >= 0 && invP[k] < neq), FEMGPU_ERR_ARG,
                 "mesh create: dof %zu is an unknown with invP_id %d out of range", k, invP[k]);
    for (size_t q = 0; q < static_cast<size_t>(nel) * sb; ++q)
        FG_CHECK(lm[q] >= 0 && lm[q] <= neq, FEMGPU_ERR_ARG, "mesh create: lm entry %d out of range", lm[q]);
    auto *M = new femgpu_mesh();
    M->ctx = ctx;
    M->ndof = ndof;
    M->nnodes = nnodes;
    M->nel = nel;
    M->neq = neq;
    M->nnz = scheme ? nnz : 0;
    hipStream_t s = ctx->stream;
    int rc = FEMGPU_OK;
    auto up = [&](auto &buf, auto *h, size_t n) {
        if (rc == FEMGPU_OK) rc = buf.upload(h, n, s);
    };
    const size_t nd = static_cast<size_t>(nnodes) * ndof;
    up(M->x, x, nnodes);
    up(M->y, y, nnodes);
    up(M->ubc, presc, nd);
    up(M->type, type, nd);
    up(M->invP, invP, nd);
    up(M->ien, ien, static_cast<size_t>(nel) * 3);
    up(M->lm, lm, static_cast<size_t>(nel) * sb);
    std::vector<std::pair<int, int>> pairs;
    pairs.reserve(static_cast<size_t>(nel) * sb);
    for (int e = 0; e < nel; ++e)
        for (int i = 0; i < sb; ++i) {
            const int I = lm[static_cast<size_t>(sb) * e + i];
            if (I < neq) pairs.emplace_back(I, sb * e + i);
        }
    std::vector<int> ptr, src, dst;
    csr_map(neq, pairs, ptr, src, dst);
    up(M->rhs_ptr, ptr.data(), ptr.size());
    up(M->rhs_src, src.data(), src.size());
    up(M->rhs_dst, dst.data(), dst.size());
    // the uploads are asynchronous: the host lists must outlive them (the aa map uses new vectors)
    std::vector<int> aptr, asrc, adst;
    if (rc == FEMGPU_OK && scheme) {
        FG_CHECK(static_cast<long long>(nel) * ss < (1LL << 31), FEMGPU_ERR_ARG, "mesh create: too many element entries");
        pairs.clear();
        for (int e = 0; e < nel; ++e)
            for (int q = 0; q < ss; ++q) {
                const int slot = scheme[static_cast<size_t>(ss) * e + q];
                if (slot < 0 || slot > nnz) rc = FEMGPU_ERR_ARG;
                else if (slot < nnz) pairs.emplace_back(slot, ss * e + q);
            }
        if (rc == FEMGPU_ERR_ARG) set_error("mesh create: scheme slot out of range");
        if (rc == FEMGPU_OK) {
            csr_map(nnz, pairs, aptr, asrc, adst);
            up(M->aa_ptr, aptr.data(), aptr.size());
            up(M->aa_src, asrc.data(), asrc.size());
            up(M->aa_dst, adst.data(), adst.size());
        }
    }
    if (rc == FEMGPU_OK) rc = M->fe.alloc(static_cast<size_t>(nel) * sb);
    if (rc == FEMGPU_OK) rc = M->ke.alloc(static_cast<size_t>(nel) * ss);
    if (rc == FEMGPU_OK && hipStreamSynchronize(s) != hipSuccess) rc = FEMGPU_ERR_HIP;
    if (rc != FEMGPU_OK) {
        delete M;
        return rc;
    }
    *out = M;
    return FEMGPU_OK;
}

int launch_block_gathers(femgpu_mesh *M, const double *Ke, double *F, double *AA)
{
    femgpu_ctx *ctx = M->ctx;
    k_rhs_gather<<<(M->neq + 1 + 255) / 256, 256, 0, ctx->stream>>>(M->neq, M->rhs_dst.p, M->rhs_ptr.p, M->rhs_src.p,
                                                                    M->fe.p, F);
    FG_HIP(hipGetLastError());
    if (AA) {
        k_aa_gather_blocks<<<(M->nnz + 1 + 255) / 256, 256, 0, ctx->stream>>>(M->nnz, M->aa_dst.p, M->aa_ptr.p,
                                                                              M->aa_src.p, Ke, AA);
        FG_HIP(hipGetLastError());
    }
    return FEMGPU_OK;
}

}  // namespace fg

extern "C" int femgpu_ns_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x, const double *y,
                                     const int *type3, const int *invP3, const double *presc3, const int *ien,
                                     const int *lm, const int *scheme, int nnz, femgpu_mesh **out)
{
    return block_mesh_create(ctx, 3, nnodes, nel, neq, x, y, type3, invP3, presc3, ien, lm, scheme, nnz, out);
}

extern "C" int femgpu_assemble_ns(femgpu_mesh *M, double Re, const double *d_u, const double *d_fx,
                                  const double *d_fy, double *d_Ke, double *d_F, double *d_AA)
{
    FG_CHECK(M && d_u && d_F && Re > 0.0, FEMGPU_ERR_ARG, "femgpu_assemble_ns: bad arguments");
    FG_CHECK(M->ndof == 3, FEMGPU_ERR_STATE, "femgpu_assemble_ns: not a Navier-Stokes mesh");
    FG_CHECK(!d_AA || M->aa_ptr.p, FEMGPU_ERR_STATE, "femgpu_assemble_ns: mesh created without a CSR slot table");
    femgpu_ctx *ctx = M->ctx;
    double *Ke = d_Ke ? d_Ke : M->ke.p;
    NsIn in;
    in.x = M->x.p;
    in.y = M->y.p;
    in.presc = M->ubc.p;
    in.u = d_u;
    in.fx = d_fx;
    in.fy = d_fy;
    in.type = M->type.p;
    in.invP = M->invP.p;
    in.ien = M->ien.p;
    in.Re = Re;
    if (M->nel > 0) {
        k_elem_ns<<<(M->nel + 255) / 256, 256, 0, ctx->stream>>>(M->nel, in, Ke, M->fe.p);
        FG_HIP(hipGetLastError());
    }
    return launch_block_gathers(M, Ke, d_F, d_AA);
}
