// Nonlinear and time drivers around the device hot path: the steady transport fixed point
// (01_SS_Transport_Equation/2D/01_CommonFiles/Process.c:29-86), the Navier-Stokes Picard loop
// (04_SS_NavierStokes_Equations/2D/01_CommonFiles/Process.c:27-52) and one step of the
// predictor-multicorrector (00_CommonFiles/Time_Integration/Predictor_Old.c:53-123 with the stop
// criteria of Criteria.c:25-41) for the transient transport and for the Euler equations (the
// latter with the EBE scheme and the nodal block-diagonal preconditioner). The state vectors stay
// in device memory; per nonlinear iteration the host reads back only the two norms the stopping
// tests need.
//
// Each linear solve is either the reference's pair precond_setup + solver (ILU(0) factorisation of
// the assembled CSR values, F := M^-1 F, pgmres on the preconditioned system: solver = 0) or the
// mixed-precision solver on the raw right-hand side (solver = 1).
#include <cmath>
#include <cstring>

#include "common.h"

using namespace fg;

namespace {

struct Work {
    double *F, *X, *uold;
};

int driver_work(femgpu_ctx *ctx, int n, Work &w)
{
    const size_t ld = (static_cast<size_t>(n) + 64) / 64 * 64;
    FG_TRY(ctx->drv_ws.alloc(3 * ld));
    w.F = ctx->drv_ws.p;
    w.X = w.F + ld;
    w.uold = w.X + ld;
    return FEMGPU_OK;
}

int check_opts(const femgpu_driver_opts *o)
{
    FG_CHECK(o && (o->solver == 0 || o->solver == 1) && o->kmax >= 2 && o->lmax >= 1 && o->tol > 0 &&
                 o->nonlinear_maxiter >= 1,
             FEMGPU_ERR_ARG, "driver: bad options");
    return FEMGPU_OK;
}

// One linear solve of the assembled system A x = F (F: raw right-hand side, overwritten by the
// reference path with M^-1 F as ILUp_precond_setup leaves it).
int linear_solve(femgpu_ctx *ctx, femgpu_csr *A, const femgpu_driver_opts *o, double *F, double *X,
                 femgpu_solve_info *si)
{
    femgpu_system S = {A, nullptr, A, nullptr};
    if (o->solver == 0) {
        FG_TRY(femgpu_ilu0_factor(A));                  // ILUp_precond_setup.c:8-61, p = 0
        FG_TRY(femgpu_ilu0_apply(A, F, F));             // F := M^-1 F (:57)
        return femgpu_pgmres(ctx, &S, o->kmax, o->lmax, o->tol, F, X, si);
    }
    if (o->mixed.factor_sweeps == 0) FG_TRY(femgpu_ilu0_factor(A));
    return femgpu_gmres_mixed(ctx, &S, &o->mixed, o->tol, F, X, si);
}

void record(femgpu_driver_info *info, int it, const femgpu_solve_info &si)
{
    if (!info) return;
    info->linear_iterations += si.iterations;
    if (it - 1 < static_cast<int>(sizeof(info->iters) / sizeof(info->iters[0]))) info->iters[it - 1] = si.iterations;
    if (si.status != 0) info->status = 2;
}

double norm(femgpu_ctx *ctx, int n, const double *x, int &rc)
{
    double v = 0.0;
    if (rc == FEMGPU_OK) rc = femgpu_ddot(ctx, n, x, x, &v);
    return std::sqrt(v);
}

}  // namespace

extern "C" int femgpu_st_process(femgpu_mesh *M, femgpu_csr *A, const femgpu_transport_problem *pb,
                                 const femgpu_driver_opts *o, double *d_u, femgpu_driver_info *info)
{
    FG_CHECK(M && A && pb && d_u && M->ndof == 1 && A->n == M->neq, FEMGPU_ERR_ARG, "femgpu_st_process: bad arguments");
    FG_TRY(check_opts(o));
    femgpu_ctx *ctx = M->ctx;
    const int neq = M->neq;
    Work w;
    FG_TRY(driver_work(ctx, neq, w));
    if (info) memset(info, 0, sizeof(*info));
    const double tol = o->nonlinear_tol;
    double normdiff = tol + 1.0, normu = 1.0 / tol;   // Process.c:29-30
    int it = 0, rc = FEMGPU_OK;
    while (normdiff > tol * normu && it < o->nonlinear_maxiter) {   // Process.c:40 (plus the cap)
        ++it;
        FG_TRY(femgpu_dcopy(ctx, neq, d_u, w.uold));
        FG_TRY(femgpu_assemble_transport(M, pb, d_u, nullptr, w.F, A->AA.p));   // Build_K_F
        FG_TRY(femgpu_csr_values_changed(A));
        femgpu_solve_info si;
        FG_TRY(linear_solve(ctx, A, o, w.F, d_u, &si));                        // solver(F, U)
        record(info, it, si);
        FG_TRY(femgpu_daxpy(ctx, neq, -1.0, d_u, w.uold));                      // uold = uold - u
        normdiff = norm(ctx, neq, w.uold, rc);
        normu = norm(ctx, neq, d_u, rc);
        FG_TRY(rc);
    }
    if (info) {
        info->iterations = it;
        info->normdiff = normdiff;
        info->normu = normu;
        if (info->status == 0 && normdiff > tol * normu) info->status = 1;
    }
    return FEMGPU_OK;
}

extern "C" int femgpu_ns_process(femgpu_mesh *M, femgpu_csr *A, double Re, const double *d_fx, const double *d_fy,
                                 const femgpu_driver_opts *o, double *d_u, femgpu_driver_info *info)
{
    FG_CHECK(M && A && d_u && M->ndof == 3 && A->n == M->neq, FEMGPU_ERR_ARG, "femgpu_ns_process: bad arguments");
    FG_TRY(check_opts(o));
    femgpu_ctx *ctx = M->ctx;
    const int neq = M->neq;
    Work w;
    FG_TRY(driver_work(ctx, neq, w));
    if (info) memset(info, 0, sizeof(*info));
    double delta2 = o->nonlinear_tol + 1.0, norms = 0.0, normu = 0.0;   // Process.c:30-31
    int it = 0, rc = FEMGPU_OK;
    while (delta2 > o->nonlinear_tol && it < o->nonlinear_maxiter) {   // Process.c:33
        ++it;
        FG_TRY(femgpu_assemble_ns(M, Re, d_u, d_fx, d_fy, nullptr, w.F, A->AA.p));   // Build
        FG_TRY(femgpu_csr_values_changed(A));
        femgpu_solve_info si;
        FG_TRY(linear_solve(ctx, A, o, w.F, w.X, &si));                           // solver(F, s)
        record(info, it, si);
        FG_TRY(femgpu_daxpy(ctx, neq, 1.0, w.X, d_u));                            // u = u + s
        norms = norm(ctx, neq, w.X, rc);
        normu = norm(ctx, neq, d_u, rc);
        FG_TRY(rc);
        delta2 = norms / normu;
    }
    if (info) {
        info->iterations = it;
        info->normdiff = norms;
        info->normu = normu;
        if (info->status == 0 && delta2 > o->nonlinear_tol) info->status = 1;
    }
    return FEMGPU_OK;
}

extern "C" int femgpu_tt_predictor_old_step(femgpu_mesh *M, femgpu_csr *A, const femgpu_transport_problem *pb,
                                            double dt, double alpha, const femgpu_driver_opts *o, double *d_u,
                                            double *d_a, femgpu_driver_info *info)
{
    FG_CHECK(M && A && pb && d_u && d_a && M->ndof == 1 && A->n == M->neq && dt > 0, FEMGPU_ERR_ARG,
             "femgpu_tt_predictor_old_step: bad arguments");
    FG_TRY(check_opts(o));
    femgpu_ctx *ctx = M->ctx;
    const int neq = M->neq;
    Work w;
    FG_TRY(driver_work(ctx, neq, w));
    if (info) memset(info, 0, sizeof(*info));
    // predictor (Predictor_Old.c:63-66)
    FG_TRY(femgpu_dcopy(ctx, neq, d_u, w.uold));
    FG_TRY(femgpu_daxpy(ctx, neq, (1.0 - alpha) * dt, d_a, d_u));
    FG_TRY(femgpu_dzero(ctx, neq, d_a));
    double norm_a = 0.0, norm_Da = 0.0;
    int i = 0, rc = FEMGPU_OK;
    for (;;) {   // multicorrection (:68-116)
        ++i;
        FG_TRY(femgpu_assemble_transport_transient(M, pb, alpha, dt, d_u, d_a, nullptr, w.F, A->AA.p));   // Build
        FG_TRY(femgpu_csr_values_changed(A));
        femgpu_solve_info si;
        FG_TRY(linear_solve(ctx, A, o, w.F, w.X, &si));   // precond_setup + solver(R, DAa)
        record(info, i, si);
        FG_TRY(femgpu_daxpy(ctx, neq, 1.0, w.X, d_a));           // a += Da
        FG_TRY(femgpu_daxpy(ctx, neq, alpha * dt, w.X, d_u));    // u += alpha dt Da
        norm_a = norm(ctx, neq, d_a, rc);
        norm_Da = norm(ctx, neq, w.X, rc);
        FG_TRY(rc);
        // StopByNorm / StopByIterations (Criteria.c:25-41)
        if (o->stop_by_norm ? (norm_Da < o->nonlinear_tol * norm_a || i >= o->nonlinear_maxiter)
                            : (i >= o->nonlinear_maxiter))
            break;
    }
    if (info) {
        info->iterations = i;
        info->norm_a = norm_a;
        info->norm_Da = norm_Da;
        if (info->status == 0 && o->stop_by_norm && !(norm_Da < o->nonlinear_tol * norm_a)) info->status = 1;
    }
    return FEMGPU_OK;
}

// Euler linear solve of one multicorrection (Predictor_Old.c:95-96 with the EBE scheme): the
// nodal block-diagonal blocks from the element matrices (BlockDiag_precond_EBE_setup), then
// F := D^-1 F and pgmres with the gathered 12x12 product and BlockDiag_precond (solver 0), or the
// mixed solver on the raw F with the FP32 copies of both (solver 1).
static int euler_solve(femgpu_ctx *ctx, femgpu_ebe *E, femgpu_blockdiag *B, const femgpu_driver_opts *o,
                       const double *Ae, double *F, double *X, femgpu_solve_info *si)
{
    femgpu_system S = {nullptr, E, nullptr, B};
    FG_TRY(femgpu_ebe_set_values(E, Ae, 1));
    FG_TRY(femgpu_blockdiag_setup(B, Ae));
    if (o->solver == 0) {
        FG_TRY(femgpu_blockdiag_apply(B, F, F));
        return femgpu_pgmres(ctx, &S, o->kmax, o->lmax, o->tol, F, X, si);
    }
    femgpu_mixed_opts mo = o->mixed;
    mo.graphs = 0;
    return femgpu_gmres_mixed(ctx, &S, &mo, o->tol, F, X, si);
}

extern "C" int femgpu_eu_predictor_old_step(femgpu_mesh *M, femgpu_ebe *E, femgpu_blockdiag *B,
                                            const femgpu_euler_problem *pb, const femgpu_driver_opts *o,
                                            double *d_u, double *d_a, femgpu_driver_info *info)
{
    FG_CHECK(M && E && B && pb && d_u && d_a && M->ndof == 4 && E->s == 12 && E->NEQ == M->neq && E->nel == M->nel &&
                 pb->dt > 0,
             FEMGPU_ERR_ARG, "femgpu_eu_predictor_old_step: bad arguments");
    FG_TRY(check_opts(o));
    FG_TRY(blockdiag_size_check(B, M->neq, M->nnodes, 4));
    femgpu_ctx *ctx = M->ctx;
    const int neq = M->neq;
    const double dt = pb->dt, alpha = pb->alpha;
    Work w;
    FG_TRY(driver_work(ctx, neq, w));
    FG_TRY(ctx->drv_ae.alloc(static_cast<size_t>(M->nel) * 144));
    if (info) memset(info, 0, sizeof(*info));
    // predictor (Predictor_Old.c:63-66)
    FG_TRY(femgpu_dcopy(ctx, neq, d_u, w.uold));
    FG_TRY(femgpu_daxpy(ctx, neq, (1.0 - alpha) * dt, d_a, d_u));
    FG_TRY(femgpu_dzero(ctx, neq, d_a));
    double norm_a = 0.0, norm_Da = 0.0;
    int i = 0, rc = FEMGPU_OK;
    for (;;) {   // multicorrection (:68-116)
        ++i;
        FG_TRY(femgpu_assemble_euler(M, pb, d_u, d_a, ctx->drv_ae.p, w.F, nullptr));   // Build_M_K_F_SUPG
        femgpu_solve_info si;
        FG_TRY(euler_solve(ctx, E, B, o, ctx->drv_ae.p, w.F, w.X, &si));               // precond_setup + solver
        record(info, i, si);
        FG_TRY(femgpu_daxpy(ctx, neq, 1.0, w.X, d_a));           // a += Da
        FG_TRY(femgpu_daxpy(ctx, neq, alpha * dt, w.X, d_u));    // u += alpha dt Da
        norm_a = norm(ctx, neq, d_a, rc);
        norm_Da = norm(ctx, neq, w.X, rc);
        FG_TRY(rc);
        if (o->stop_by_norm ? (norm_Da < o->nonlinear_tol * norm_a || i >= o->nonlinear_maxiter)
                            : (i >= o->nonlinear_maxiter))
            break;
    }
    if (info) {
        info->iterations = i;
        info->norm_a = norm_a;
        info->norm_Da = norm_Da;
        if (info->status == 0 && o->stop_by_norm && !(norm_Da < o->nonlinear_tol * norm_a)) info->status = 1;
    }
    return FEMGPU_OK;
}
