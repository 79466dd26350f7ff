/*
 * ref_harness_st.c -- one-rank driver around the UNMODIFIED reference hot-path sources of
 * 01_SS_Transport_Equation (compiled in place from the reference tree by oracle/ref.mk).
 * Test infrastructure: it writes golden vectors for tests/golden and is never shipped. Its shim
 * build (below) is the one form that runs on a GPU machine: the test of the reference-side binding.
 *
 * It plays the role of Preprocess.c:127-357 and Process.c:19-86 for NPROC = 1 (the real
 * drivers refuse fewer than two ranks, Preprocess.c:50-53) and supplies:
 *   - the problem callbacks (Condutivity/Font/Reaction/Velocity/upresc) of the build's own
 *     benchmark problems, data-driven from the input file;
 *   - MPI_VectorUpdate for one rank: there is no neighbour, so there is nothing to exchange
 *     (the driver's own file hard-codes a send to rank 1, MPI_VectorUpdate.c:25-29);
 *   - a ShockCapture callback returning 0 for the linear mode (Eu = 0).
 * Everything on the path -- Fill_LM, Permutation_of_LM, csr_Initialization, reordering,
 * symrcm, Build_K_F_SUPG, F_assembly, CAU_ShockCapture, h_shock_2sqrtArea, ebe/csr_assembly,
 * setzeros, ebemv, csrmv, ILUp_precond(_setup), pgmres and the BLAS -- is the reference's code.
 *
 * Input (binary, little endian): int32 header[16], float64 params[16], then x[nn], y[nn],
 * type[nn] (int32), ien[3*nel] (int32), ubc[nn], probe[nn] (float64).
 * Built with -DFEMGPU_SHIM and linked with parallel_fem_codes_amd/compat/femgpu_compat.c and
 * libfemgpu.so in place of the reference's ebemv, csrmv, ILUp_precond(_setup), NO_precond_setup
 * and pgmres objects, it runs the reference's own assembly around the library's device solver.
 * Output: a stream of records {char name[32]; int32 kind (0 int32, 1 float64); int64 count;
 * payload}. The record "timers" holds the wall seconds spent in Build_K_F_SUPG, precond_setup and
 * pgmres, summed over the nonlinear iterations (clock_gettime, CLOCK_MONOTONIC).
 */
#include <time.h>
#include "SSTranspEquation.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/preconditioners.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/solvers.h"
#include "../../../00_CommonFiles/MatrixVector_Operations/matvec.h"
#include "../../../00_CommonFiles/BLAS_Operations/ourBLAS.h"

static double g_k, g_gamma, g_f0, g_bx[3], g_by[3];
static int g_nn;
static const double *g_x, *g_y, *g_ubc;

static double h_cond(void) { return g_k; }
static double h_react(void) { return g_gamma; }
static double h_font(double X, double Y, double k, double gamma, double bx, double by) { return g_f0; }
static void h_vel(double X, double Y, double B[2])
{
    B[0] = g_bx[0] + g_bx[1] * X + g_bx[2] * Y;
    B[1] = g_by[0] + g_by[1] * X + g_by[2] * Y;
}
/* Dirichlet data is given per node; the reference evaluates upresc at node coordinates, so the
 * lookup by exact coordinates returns the value of the node the reference is asking about. */
static double h_upresc(double X, double Y)
{
    for (int i = 0; i < g_nn; i++)
        if (g_x[i] == X && g_y[i] == Y) return g_ubc[i];
    return 0.0;
}
static double h_noshock(double *a, double *b, int e, double kx, double ky, double bx, double by,
                        double g, double u1, double u2, double u3, double ub, double fb, double y23,
                        double y31, double y12, double x32, double x13, double x21, double ia,
                        double h) { return 0.0; }

void MPI_VectorUpdate(int RANK, int NPROC, ParametersType *Parameters, FemStructsType *FemStructs,
                      double *P, int tag) { (void)RANK; (void)NPROC; (void)Parameters; (void)FemStructs; (void)P; (void)tag; }

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static FILE *g_out;
static void rec(const char *name, int kind, long n, const void *p)
{
    char nm[32] = {0};
    strncpy(nm, name, 31);
    fwrite(nm, 1, 32, g_out);
    int k = kind;
    long long c = n;
    fwrite(&k, sizeof(int), 1, g_out);
    fwrite(&c, sizeof(long long), 1, g_out);
    fwrite(p, kind ? sizeof(double) : sizeof(int), (size_t)n, g_out);
}

int main(int argc, char **argv)
{
    int RANK, NPROC, I, J;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &RANK);
    MPI_Comm_size(MPI_COMM_WORLD, &NPROC);
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); MPI_Abort(MPI_COMM_WORLD, 2); }
    FILE *in = fopen(argv[1], "rb");
    int hd[16];
    double pr[16];
    if (!in || fread(hd, sizeof(int), 16, in) != 16 || fread(pr, sizeof(double), 16, in) != 16) MPI_Abort(MPI_COMM_WORLD, 3);
    const int nn = hd[0], nel = hd[1], ebe = hd[2], reorder = hd[3], shock = hd[4], kmax = hd[5],
              lmax = hd[6], nl_max = hd[7];
    g_k = pr[0]; g_gamma = pr[1]; g_f0 = pr[2];
    for (I = 0; I < 3; I++) { g_bx[I] = pr[3 + I]; g_by[I] = pr[6 + I]; }
    const double tol = pr[9], nltol = pr[10];
    double *x = calloc(nn, sizeof(double)), *y = calloc(nn, sizeof(double));
    double *ubc = calloc(nn, sizeof(double)), *probe = calloc(nn, sizeof(double));
    int *type = calloc(nn, sizeof(int)), *ien = calloc(3 * (size_t)nel, sizeof(int));
    size_t ok = fread(x, sizeof(double), nn, in) + fread(y, sizeof(double), nn, in) +
                fread(type, sizeof(int), nn, in) + fread(ien, sizeof(int), 3 * (size_t)nel, in) +
                fread(ubc, sizeof(double), nn, in) + fread(probe, sizeof(double), nn, in);
    fclose(in);
    if (ok != (size_t)5 * nn + 3 * (size_t)nel) MPI_Abort(MPI_COMM_WORLD, 4);
    g_nn = nn; g_x = x; g_y = y; g_ubc = ubc;
    g_out = fopen(argv[2], "wb");

    /* ---- Preprocess for one rank (Preprocess.c:127-357 with NPROC = 1, no halo lists). ---- */
    ParametersType *P = calloc(1, sizeof(ParametersType));
    MatrixDataType *MD = calloc(1, sizeof(MatrixDataType));
    FemStructsType *FS = calloc(1, sizeof(FemStructsType));
    FemFunctionsType *FF = calloc(1, sizeof(FemFunctionsType));
    NodeType *Node = calloc(nn, sizeof(NodeType));
    ElementType *Element = calloc(nel, sizeof(ElementType));
    int neq = 0;
    for (I = 0; I < nn; I++) {
        Node[I].x = x[I]; Node[I].y = y[I]; Node[I].Type = type[I]; Node[I].Send = 1;
        Node[I].invP_id = (type[I] == 0) ? -1 : neq++;
    }
    for (I = 0; I < nel; I++)
        for (J = 0; J < 3; J++) Element[I].Vertex[J] = ien[3 * I + J];
    P->neq = neq; P->neq_bef = 0; P->NEQ = neq; P->nel = nel; P->nnodes = nn; P->RANK = 0; P->NPROC = 1;
    P->KrylovBasisVectorsQuantity = kmax; P->LinearMaxIter = lmax; P->SolverTolerance = tol;
    P->NonLinearTolerance = nltol;
    strcpy(P->Preconditioner, ebe ? "NOT" : "ILU0");
    strcpy(P->reordering, reorder ? "SYMRCM" : "NOT");
    strcpy(P->MatrixVectorProductScheme, ebe ? "EBE" : "CSR");
    double *F = calloc(neq + 1, sizeof(double)), *U = calloc(neq + 1, sizeof(double));
    int **lm = calloc(nel, sizeof(int *)), *lmaux = calloc(3 * (size_t)nel, sizeof(int));
    for (I = 0; I < nel; I++) lm[I] = &lmaux[3 * I];
    Fill_LM(P, lm, Node, Element);
    FS->Node = Node; FS->Element = Element; FS->lm = lm; FS->lmaux = lmaux;
    FS->F = F; FS->U = U; FS->f = F; FS->u = U;
    MD->Diag = calloc(neq + 1, sizeof(double)); MD->invDiag = calloc(neq + 1, sizeof(double));
    rec("lm", 0, 3L * nel, lmaux);
    if (ebe) {
        int *Perm = calloc(neq + 1, sizeof(int)), *invPerm = calloc(neq + 1, sizeof(int));
        Permutation_of_LM(P, FS, Perm, invPerm);
        double **K = calloc(nel, sizeof(double *)), *Kaux = calloc(9 * (size_t)nel, sizeof(double));
        for (I = 0; I < nel; I++) K[I] = &Kaux[9 * I];
        MD->A = K; MD->Aaux = Kaux; MD->Perm = Perm; MD->invPerm = invPerm;
        rec("Perm", 0, neq + 1, Perm);
        rec("invPerm", 0, neq + 1, invPerm);
        FF->assembly = ebe_assembly; FF->mv = ebemv;
        FF->precond = NO_precond; FF->precondR = NO_precond; FF->precond_setup = NO_precond_setup;
    } else {
        csr_Initialization(P, MD, FS);
        MD->AA = calloc(P->nnzero + 1, sizeof(double));
        MD->AA_bef = calloc(P->nnzero_bef + 1, sizeof(double));
        MD->AA_aft = calloc(P->nnzero_aft + 1, sizeof(double));
        rec("IA", 0, neq + 1, MD->IA);
        rec("JA", 0, P->nnzero, MD->JA);
        rec("perm", 0, neq, MD->Perm);
        int *sch = calloc(9 * (size_t)nel, sizeof(int));
        for (I = 0; I < nel; I++)
            for (J = 0; J < 9; J++) sch[9 * I + J] = MD->Scheme_by_Element[I][J];
        rec("scheme", 0, 9L * nel, sch);
        free(sch);
        FF->assembly = csr_assembly; FF->mv = csrmv;
        FF->precond = ILUp_precond; FF->precondR = NO_precond; FF->precond_setup = ILUp_precond_setup;
    }
    int *invp = calloc(nn, sizeof(int));
    for (I = 0; I < nn; I++) invp[I] = Node[I].invP_id;
    rec("invP_id", 0, nn, invp);
    FF->Condutivity = h_cond; FF->Font = h_font; FF->Reaction = h_react; FF->Velocity = h_vel;
    FF->upresc = h_upresc; FF->h_shock_calculation = h_shock_2sqrtArea;
    FF->ShockCapture = shock ? CAU_ShockCapture : h_noshock;
    FS->delta_old = calloc(nel, sizeof(double)); FS->Re_old = calloc(nel, sizeof(double));

    /* ---- Process.c:29-86 loop with the reference Build / precond_setup / pgmres. ---- */
    double *uold = calloc(neq + 1, sizeof(double));
    double normdiff = nltol + 1, normu = 1.0 / nltol;
    int it = 0;
    int *iters = calloc(nl_max + 1, sizeof(int));
    double timers[3] = {0, 0, 0}, t0;
    while (normdiff > nltol * normu && it < nl_max) {
        it++;
        dcopy(neq, U, uold);
        t0 = now();
        Build_K_F_SUPG(P, MD, FS, FF);
        timers[0] += now() - t0;
        if (it == 1) {
            if (ebe) rec("A", 1, 9L * nel, MD->Aaux); else rec("AA", 1, P->nnzero + 1, MD->AA);
            rec("F", 1, neq, F);
            /* operator probe on the first matrix: probe vector given per node, read at unknowns */
            double *pv = calloc(neq + 1, sizeof(double)), *qv = calloc(neq + 1, sizeof(double));
            for (I = 0; I < nn; I++) if (Node[I].Type == 1) pv[Node[I].invP_id] = probe[I];
            FF->mv(P, MD, FS, pv, qv);
            rec("mv", 1, neq, qv);
            free(pv); free(qv);
        }
        t0 = now();
        FF->precond_setup(P, MD, FS, it, F);
        timers[1] += now() - t0;
        if (it == 1) {
            rec("MinvF", 1, neq, F);
#ifndef FEMGPU_SHIM   /* with the library's shim the factors live on the device, not in MD->ILUp */
            if (!ebe) {
                SparILU *lu = MD->ILUp;
                long nl = 0, nu = 0;
                for (I = 0; I < neq; I++) { nl += lu->L->nzcount[I]; nu += lu->U->nzcount[I]; }
                double *Lv = calloc(nl + 1, sizeof(double)), *Uv = calloc(nu + 1, sizeof(double));
                long a = 0, b = 0;
                for (I = 0; I < neq; I++) {
                    for (J = 0; J < lu->L->nzcount[I]; J++) Lv[a++] = lu->L->ma[I][J];
                    for (J = 0; J < lu->U->nzcount[I]; J++) Uv[b++] = lu->U->ma[I][J];
                }
                rec("ILU_L", 1, nl, Lv); rec("ILU_U", 1, nu, Uv); rec("ILU_D", 1, neq, lu->D);
                free(Lv); free(Uv);
            }
#endif
        }
        const int before = P->iterations;
        t0 = now();
        pgmres(P, MD, FS, FF, F, U);
        timers[2] += now() - t0;
        iters[it - 1] = P->iterations - before;
        if (it == 1) rec("X1", 1, neq, U);
        daxpy(neq, -1, U, uold);
        normdiff = sqrt(ddot(neq, uold, uold));
        normu = sqrt(ddot(neq, U, U));
    }
    rec("iters", 0, it, iters);
    rec("timers", 1, 3, timers);
    rec("u", 1, neq, U);
    fclose(g_out);
    MPI_Finalize();
    return 0;
}
