// Per-element SUPG assembly of the 2D compressible Euler equations (P1, conservative variables
// rho, rho u, rho v, rho E; 12x12 element blocks) on the device: Ae = Me + alpha*dt*Ke and the
// residual R = -(Me dU + Ke U) of Build_M_K_F_SUPG.c:6-1057, with eval_U_dU.c:3-55, the flux
// Jacobians of Ax_Ay_calculations.c (dimensionless :3-53 or dimensional :55-99) and the CAU
// shock-capture coefficient of Delta_CAU.c:4-83.
//
// One thread per element. The element matrices are built row by row: for dof row i of vertex a,
// the Galerkin/SUPG convection block C_a = (b_a Ax + c_a Ay)/6, the SUPG products
// b_a (AxAx, AxAy) + c_a (AyAx, AyAy) and the CAU coupling give every column, and the row's
// residual and Ae entries are formed on the spot. Each quantity is evaluated in the reference's
// operation order with contraction off; the reference's hand-expanded 4x4 products skip only
// terms with a structural zero factor, so the dense products used here have equal values.
// Global accumulation is the element-order gather shared with the Navier-Stokes assembly.
#include "common.h"

namespace fg {

struct EulerIn {
    const double *x, *y, *presc, *u, *du;
    const int *type, *invP, *ien;
    femgpu_euler_problem pb;
};

__device__ __forceinline__ void euler_jacobians_dev(const femgpu_euler_problem &p, const double Ub[4],
                                                    double Ax[4][4], double Ay[4][4])
{
#pragma clang fp contract(off)
    const double g = p.gamma;
    const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Ax[i][j] = 0.0;
            Ay[i][j] = 0.0;
        }
    Ax[0][1] = 1.0;
    Ay[0][2] = 1.0;
    Ax[2][0] = -v1 * v2;
    Ax[2][1] = v2;
    Ax[2][2] = v1;
    Ax[1][1] = (3.0 - g) * v1;
    Ax[1][2] = -(g - 1.0) * v2;
    Ax[3][3] = g * v1;
    Ay[1][1] = v2;
    Ay[1][2] = v1;
    Ay[2][1] = -(g - 1.0) * v1;
    Ay[2][2] = (3.0 - g) * v2;
    Ay[3][3] = g * v2;
    if (p.dimensionless) {
        const double nv = v1 * v1 + v2 * v2, eps2 = p.Mach * p.Mach;
        Ax[1][0] = -(v1 * v1) + ((g - 1) * nv * 0.5);
        Ax[1][3] = (g - 1.0) / eps2;
        Ax[3][0] = -v1 * (g * E - ((g - 1.0) * eps2 * nv));
        Ax[3][1] = g * E - (g - 1.0) * eps2 * (0.5 * nv + v1 * v1);
        Ax[3][2] = -(g - 1.0) * eps2 * v1 * v2;
        Ay[2][0] = -(v2 * v2) + (((g - 1.0) * nv * 0.5));
        Ay[3][0] = -v2 * (g * E - (g - 1.0) * eps2 * nv);
        Ay[3][2] = g * E - (g - 1.0) * eps2 * (0.5 * nv + v2 * v2);
    } else {
        const double rho = Ub[0], nU = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        Ax[1][0] = -(v1 * v1) + (((g - 1) * nU * 0.5) / (rho * rho));
        Ax[1][3] = g - 1.0;
        Ax[3][0] = (-g * v1 * E) + ((g - 1.0) * v1 * (nU / (rho * rho)));
        Ax[3][1] = (g * E) - ((g - 1.0) * 0.5 * ((nU / (rho * rho)) + (2.0 * v1 * v1)));
        Ax[3][2] = -((g - 1.0) * v1 * v2);
        Ay[2][0] = -(v2 * v2) + (((g - 1.0) * nU * 0.5) / (rho * rho));
        Ay[3][0] = (-g * v2 * E) + (((g - 1.0) * v2 * nU) / (rho * rho));
        Ay[3][2] = (g * E) - ((g - 1.0) * 0.5 * ((nU / (rho * rho)) + (2.0 * v2 * v2)));
    }
    Ay[1][0] = Ax[2][0];
    Ay[2][3] = Ax[1][3];
    Ay[3][1] = Ax[3][2];
}

// w^T A0 w grouped as Delta_CAU.c:39-73 groups it.
__device__ __forceinline__ double a0_form_dev(const double A0[4][4], const double w[4])
{
#pragma clang fp contract(off)
    return w[0] * (A0[0][0] * w[0] + 2.0 * A0[0][1] * w[1] + 2.0 * A0[0][2] * w[2] + 2.0 * A0[0][3] * w[3]) +
           w[1] * (A0[1][1] * w[1] + 2.0 * A0[1][2] * w[2] + 2.0 * A0[1][3] * w[3]) +
           w[2] * (A0[2][2] * w[2] + 2.0 * A0[2][3] * w[3]) + w[3] * A0[3][3] * w[3];
}

// Entry (i, j) of A B, accumulated over k from the k = 0 term.
__device__ __forceinline__ double mul4_entry(const double A[4][4], const double B[4][4], int i, int j)
{
#pragma clang fp contract(off)
    double acc = A[i][0] * B[0][j];
#pragma unroll
    for (int k = 1; k < 4; ++k) acc += A[i][k] * B[k][j];
    return acc;
}

__global__ __launch_bounds__(256) void k_elem_euler(int nel, EulerIn in, double *__restrict__ Ae, double *__restrict__ Res)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nel) return;
    const femgpu_euler_problem &p = in.pb;
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    const double gamma = p.gamma, alpha = p.alpha, delta_t = p.dt;
    int v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = in.ien[3 * e + a];
    const double X0 = in.x[v[0]], X1 = in.x[v[1]], X2 = in.x[v[2]];
    const double Y0 = in.y[v[0]], Y1 = in.y[v[1]], Y2 = in.y[v[2]];
    const double b[3] = {Y1 - Y2, Y2 - Y0, Y0 - Y1};
    const double c[3] = {X2 - X1, X0 - X2, X1 - X0};
    const double twoArea = fabs(c[2] * b[1] - c[1] * b[2]);
    const double Area = 0.5 * twoArea;
    double Ue[12], dUe[12];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int q = 4 * v[a] + k;
            const bool unk = in.type[q] > 0;
            Ue[4 * a + k] = unk ? in.u[in.invP[q]] : in.presc[q];
            dUe[4 * a + k] = unk ? in.du[in.invP[q]] : 0.0;
        }
    double Ub[4], dUb[4], gUx[4], gUy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        Ub[i] = (Ue[i] + Ue[i + 4] + Ue[i + 8]) * third;
        dUb[i] = (dUe[i] + dUe[i + 4] + dUe[i + 8]) * third;
        gUx[i] = (Ue[i] * b[0] + Ue[i + 4] * b[1] + Ue[i + 8] * b[2]) / twoArea;
        gUy[i] = (Ue[i] * c[0] + Ue[i + 4] * c[1] + Ue[i + 8] * c[2]) / twoArea;
    }
    const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
    double Ax[4][4], Ay[4][4];
    euler_jacobians_dev(p, Ub, Ax, Ay);
    // CAU coefficient (Delta_CAU.c) with the entropy-variable metric A0^-1 (:134-160)
    double delta = 0.0;
    if (p.shock) {
        const double nU = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        const double rho_i = Ub[3] - ((nU * 0.5) / Ub[0]);
        const double V2 = Ub[1] / rho_i, V3 = Ub[2] / rho_i, V4 = (-Ub[0]) / rho_i;
        const double rhoi_e = -1.0 / (rho_i * V4);
        const double k1 = (V2 * V2 + V3 * V3) * 0.5 / V4;
        double A0[4][4];
        A0[0][0] = rhoi_e * (k1 * k1 + gamma);
        A0[0][1] = rhoi_e * (k1 * V2);
        A0[0][2] = rhoi_e * (k1 * V3);
        A0[0][3] = rhoi_e * ((k1 + 1.0) * V4);
        A0[1][1] = rhoi_e * (V2 * V2 - V4);
        A0[1][2] = rhoi_e * (V2 * V3);
        A0[1][3] = rhoi_e * (V2 * V4);
        A0[2][2] = rhoi_e * (V3 * V3 - V4);
        A0[2][3] = rhoi_e * (V3 * V4);
        A0[3][3] = rhoi_e * V4 * V4;
        A0[1][0] = A0[0][1];
        A0[2][0] = A0[0][2];
        A0[2][1] = A0[1][2];
        A0[3][0] = A0[0][3];
        A0[3][1] = A0[1][3];
        A0[3][2] = A0[2][3];
        double AgUx[4], AgUy[4], gkx[4], gky[4], LUh[4];
        AgUx[0] = gUx[1];
        AgUx[1] = Ax[1][0] * gUx[0] + Ax[1][1] * gUx[1] + Ax[1][2] * gUx[2] + Ax[1][3] * gUx[3];
        AgUx[2] = Ax[2][0] * gUx[0] + Ax[2][1] * gUx[1] + Ax[2][2] * gUx[2];
        AgUx[3] = Ax[3][0] * gUx[0] + Ax[3][1] * gUx[1] + Ax[3][2] * gUx[2] + Ax[3][3] * gUx[3];
        AgUy[0] = gUy[2];
        AgUy[1] = Ay[1][0] * gUy[0] + Ay[1][1] * gUy[1] + Ay[1][2] * gUy[2];
        AgUy[2] = Ay[2][0] * gUy[0] + Ay[2][1] * gUy[1] + Ay[2][2] * gUy[2] + Ay[2][3] * gUy[3];
        AgUy[3] = Ay[3][0] * gUy[0] + Ay[3][1] * gUy[1] + Ay[3][2] * gUy[2] + Ay[3][3] * gUy[3];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            gkx[i] = (b[0] * gUx[i] + c[0] * gUy[i]) / twoArea;
            gky[i] = (b[1] * gUx[i] + c[1] * gUy[i]) / twoArea;
            LUh[i] = dUb[i] + AgUx[i] + AgUy[i];
        }
        const double n_gk = sqrt(a0_form_dev(A0, gkx) + a0_form_dev(A0, gky));
        const double n_LUh = sqrt(a0_form_dev(A0, LUh));
        const double n_gU = sqrt(a0_form_dev(A0, gUx) + a0_form_dev(A0, gUy));
        delta = (n_gU >= p.tolerance) ? n_LUh / n_gk : 0.0;
    }
    // SUPG parameter (:246-271)
    double betax = 2.0 * (gUx[0] * Ub[0] + gUx[1] * Ub[1] + gUx[2] * Ub[2] + gUx[3] * Ub[3]);
    double betay = 2.0 * (gUy[0] * Ub[0] + gUy[1] * Ub[1] + gUy[2] * Ub[2] + gUy[3] * Ub[3]);
    const double betaxy = sqrt(betax * betax + betay * betay);
    if (betaxy != 0.0) {
        betax = betax / betaxy;
        betay = betay / betaxy;
    }
    const double h = sqrt(twoArea);
    const double abs_vbeta = fabs(v1 * betax + v2 * betay);
    const double cs = sqrt(gamma * (gamma - 1.0) * (E - 0.5 * (v1 * v1 + v2 * v2)));
    const double CFL = ((cs + abs_vbeta) * delta_t) / h;
    const double psi = (2.0 * alpha * CFL) * (1.0 / (1.0 + 2.0 * alpha * CFL));
    const double tau_d = delta / ((cs + abs_vbeta) * (cs + abs_vbeta));
    const double tau_a = h / (2.0 * (cs + abs_vbeta));
    const double tau_t = (2.0 / (3.0 * (1.0 + 2.0 * alpha * CFL))) * tau_a;
    double tau = tau_t + psi * (tau_a - tau_d);
    if (tau < 0.0) tau = 0.0;
    const double Mg1 = Area * sixth, Mg2 = Area / 12.0;
    const double tau4area = tau / (4 * Area), delta4area = delta / (4.0 * Area);
    const double Kc01 = delta4area * (b[0] * b[1] + c[0] * c[1]);
    const double Kc02 = delta4area * (b[0] * b[2] + c[0] * c[2]);
    const double Kc12 = delta4area * (b[1] * b[2] + c[1] * c[2]);
    double *A = Ae + 144 * static_cast<size_t>(e);
    double *R = Res + 12 * static_cast<size_t>(e);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // row i of the four Jacobian products, shared by the three vertex rows 4a+i
        double Pxx[4], Pxy[4], Pyx[4], Pyy[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Pxx[j] = mul4_entry(Ax, Ax, i, j);
            Pxy[j] = mul4_entry(Ax, Ay, i, j);
            Pyx[j] = mul4_entry(Ay, Ax, i, j);
            Pyy[j] = mul4_entry(Ay, Ay, i, j);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
            const double Kca1 = (a == 0) ? Kc01 : (a == 1 ? Kc01 : Kc02);   // Kcau(a, o1)
            const double Kca2 = (a == 0) ? Kc02 : Kc12;                      // Kcau(a, o2)
            double MdU = 0, KU = 0;
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
                const double Kab = (a == bb) ? 0.0 : ((a + bb == 1) ? Kc01 : (a + bb == 2 ? Kc02 : Kc12));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double Ca_ij = (b[a] * Ax[i][j] + c[a] * Ay[i][j]) * sixth;
                    const double Mpg = Ca_ij * tau;
                    const double AXij = b[a] * Pxx[j] + c[a] * Pyx[j];
                    const double AYij = b[a] * Pxy[j] + c[a] * Pyy[j];
                    double m, k;
                    if (i == j) m = (a == bb ? Mg1 : Mg2) + Mpg;
                    else m = Mpg;
                    if (a == bb) {
                        const double k1v = tau4area * (b[o1] * AXij + c[o1] * AYij);
                        const double k2v = tau4area * (b[o2] * AXij + c[o2] * AYij);
                        k = Ca_ij - (k1v + k2v);
                        if (i == j) k = k - (Kca1 + Kca2);
                    } else {
                        const double kp = tau4area * (b[bb] * AXij + c[bb] * AYij);
                        const double Cb_ij = (b[bb] * Ax[i][j] + c[bb] * Ay[i][j]) * sixth;
                        k = Cb_ij + kp;
                        if (i == j) k = k + Kab;
                    }
                    const int col = 4 * bb + j;
                    MdU += m * dUe[col];
                    KU += k * Ue[col];
                    A[12 * (4 * a + i) + col] = m + alpha * delta_t * k;
                }
            }
            R[4 * a + i] = -MdU - KU;
        }
    }
}

}  // namespace fg

using namespace fg;

extern "C" int femgpu_euler_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x,
                                        const double *y, const int *type4, const int *invP4, const double *presc4,
                                        const int *ien, const int *lm, const int *scheme, int nnz, femgpu_mesh **out)
{
    return block_mesh_create(ctx, 4, nnodes, nel, neq, x, y, type4, invP4, presc4, ien, lm, scheme, nnz, out);
}

extern "C" int femgpu_assemble_euler(femgpu_mesh *M, const femgpu_euler_problem *pb, const double *d_u,
                                     const double *d_du, double *d_Ae, double *d_R, double *d_AA)
{
    FG_CHECK(M && pb && d_u && d_du && d_R, FEMGPU_ERR_ARG, "femgpu_assemble_euler: bad arguments");
    FG_CHECK(M->ndof == 4, FEMGPU_ERR_STATE, "femgpu_assemble_euler: not an Euler mesh");
    FG_CHECK(!d_AA || M->aa_ptr.p, FEMGPU_ERR_STATE, "femgpu_assemble_euler: mesh created without a CSR slot table");
    femgpu_ctx *ctx = M->ctx;
    double *Ae = d_Ae ? d_Ae : M->ke.p;
    EulerIn in;
    in.x = M->x.p;
    in.y = M->y.p;
    in.presc = M->ubc.p;
    in.u = d_u;
    in.du = d_du;
    in.type = M->type.p;
    in.invP = M->invP.p;
    in.ien = M->ien.p;
    in.pb = *pb;
    if (M->nel > 0) {
        k_elem_euler<<<(M->nel + 255) / 256, 256, 0, ctx->stream>>>(M->nel, in, Ae, M->fe.p);
        FG_HIP(hipGetLastError());
    }
    return launch_block_gathers(M, Ae, d_R, d_AA);
}
