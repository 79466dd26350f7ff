/*
 * ref_harness_eu.c -- one-rank driver around the UNMODIFIED reference Euler assembly sources of
 * 03_Euler_Equations (compiled in place from the reference tree by oracle/ref.mk). Test
 * infrastructure: it writes golden vectors for tests/golden; it is never shipped or run on a GPU
 * machine.
 *
 * It plays the role of Preprocess.c for NPROC = 1 and calls, once, the reference's
 * Build_M_K_F_SUPG (with eval_U_dU, the reference's Ax_Ay_calculations and Delta_CAU, setzeros,
 * Fill_LM, csr_Initialization and ebe/csr_assembly) on a given state and increment. It supplies
 * the problem callbacks (gamma, rhopresc, v1presc, v2presc, epresc) and MPI_VectorUpdate for one
 * rank; prescribed values are per-node tables looked up by exact coordinates.
 *
 * With the EBE scheme and header[6] = kmax > 0 it then runs the linear solve of the Euler
 * predictor (Predictor_Old.c:95-96): the reference's BlockDiag_precond_EBE_setup (nodal blocks,
 * inverses, R := D^-1 R) and pgmres with ebemvNDOF4 and BlockDiag_precond.
 *
 * Input (binary, little endian): int32 header[16] {nn, nel, ebe, reorder, dimensionless, shock,
 * kmax, lmax}, float64 params[16] {gamma, Mach, alpha, dt, tolerance, solver tolerance}, then
 * x[nn], y[nn], type[4nn] (int32),
 * ien[3nel] (int32), presc[4nn], state[4nn], incr[4nn] (float64; values of the unknown dofs).
 * Output: records {char name[32]; int32 kind (0 int32, 1 float64); int64 count; payload}.
 */
#include "EulerEquations.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/preconditioners.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/solvers.h"
#include "../../../00_CommonFiles/MatrixVector_Operations/matvec.h"

int Fill_LM(ParametersType *, int **, NodeType *, ElementType *);
void csr_Initialization(ParametersType *, MatrixDataType *, FemStructsType *);
void ebe_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[12]);
void csr_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[12]);

static int g_nn;
static double g_gamma;
static const double *g_x, *g_y, *g_presc;

static int node_at(double X, double Y)
{
    for (int i = 0; i < g_nn; i++)
        if (g_x[i] == X && g_y[i] == Y) return i;
    return -1;
}
static double h_gamma(double X, double Y) { (void)X; (void)Y; return g_gamma; }
static double h_rho(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[4 * i]; }
static double h_v1(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[4 * i + 1]; }
static double h_v2(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[4 * i + 2]; }
static double h_e(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[4 * i + 3]; }
static double h_noshock(double t, double *d, double *gx, double *gy, double (*Ax)[4], double (*Ay)[4],
                        double (*A0)[4], double *dUb, double y23, double y31, double y12, double x32,
                        double x13, double x21, double twoA, int e, double *invY, double *Ub) { return 0.0; }

void MPI_VectorUpdate(int RANK, int NPROC, ParametersType *Parameters, FemStructsType *FemStructs,
                      double *P, int tag) { (void)RANK; (void)NPROC; (void)Parameters; (void)FemStructs; (void)P; (void)tag; }

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
    int I, J, d;
    MPI_Init(&argc, &argv);
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); MPI_Abort(MPI_COMM_WORLD, 2); }
    FILE *in = fopen(argv[1], "rb");
    int hd[16];
    double pr[16];
    if (!in || fread(hd, sizeof(int), 16, in) != 16 || fread(pr, sizeof(double), 16, in) != 16) MPI_Abort(MPI_COMM_WORLD, 3);
    const int nn = hd[0], nel = hd[1], ebe = hd[2], reorder = hd[3], dimless = hd[4], shock = hd[5];
    double *x = calloc(nn, sizeof(double)), *y = calloc(nn, sizeof(double));
    double *presc = calloc(4 * (size_t)nn, sizeof(double)), *state = calloc(4 * (size_t)nn, sizeof(double));
    double *incr = calloc(4 * (size_t)nn, sizeof(double));
    int *type = calloc(4 * (size_t)nn, sizeof(int)), *ien = calloc(3 * (size_t)nel, sizeof(int));
    size_t ok = fread(x, sizeof(double), nn, in) + fread(y, sizeof(double), nn, in) +
                fread(type, sizeof(int), 4 * (size_t)nn, in) + fread(ien, sizeof(int), 3 * (size_t)nel, in) +
                fread(presc, sizeof(double), 4 * (size_t)nn, in) + fread(state, sizeof(double), 4 * (size_t)nn, in) +
                fread(incr, sizeof(double), 4 * (size_t)nn, in);
    fclose(in);
    if (ok != (size_t)18 * nn + 3 * (size_t)nel) MPI_Abort(MPI_COMM_WORLD, 4);
    g_nn = nn; g_x = x; g_y = y; g_presc = presc; g_gamma = pr[0];
    g_out = fopen(argv[2], "wb");

    ParametersType *P = calloc(1, sizeof(ParametersType));
    MatrixDataType *MD = calloc(1, sizeof(MatrixDataType));
    FemStructsType *FS = calloc(1, sizeof(FemStructsType));
    FemFunctionsType *FF = calloc(1, sizeof(FemFunctionsType));
    NodeType *Node = calloc(nn, sizeof(NodeType));
    ElementType *Element = calloc(nel, sizeof(ElementType));
    int neq = 0;
    for (I = 0; I < nn; I++) {
        Node[I].x = x[I]; Node[I].y = y[I];
        for (d = 0; d < 4; d++) {
            Node[I].Type[d] = type[4 * I + d];
            Node[I].Send[d] = 1;
            Node[I].invP_id[d] = (type[4 * I + d] == 0) ? -1 : neq++;
        }
    }
    for (I = 0; I < nel; I++)
        for (J = 0; J < 3; J++) Element[I].Vertex[J] = ien[3 * I + J];
    P->neq = neq; P->neq_bef = 0; P->NEQ = neq; P->nel = nel; P->nnodes = nn; P->RANK = 0; P->NPROC = 1;
    P->Mach = pr[1]; P->Alpha_Build = pr[2]; P->DeltaT_Build = pr[3];
    strcpy(P->reordering, reorder ? "SYMRCM" : "NOT");
    strcpy(P->MatrixVectorProductScheme, ebe ? "EBE" : "CSR");
    double *R = calloc(neq + 1, sizeof(double)), *u = calloc(neq + 1, sizeof(double));
    double *du = calloc(neq + 1, sizeof(double));
    int **lm = calloc(nel, sizeof(int *)), *lmaux = calloc(12 * (size_t)nel, sizeof(int));
    for (I = 0; I < nel; I++) lm[I] = &lmaux[12 * I];
    Fill_LM(P, lm, Node, Element);
    FS->Node = Node; FS->Element = Element; FS->lm = lm; FS->lmaux = lmaux;
    FS->F = R; FS->u = u; FS->du = du;
    FS->delta_old = calloc(nel, sizeof(double));
    FS->AuxBuild = calloc(1, sizeof(AuxBuildStructuresType));
    FS->AuxBuild->tolerance = pr[4];
    rec("lm", 0, 12L * nel, lmaux);
    if (ebe) {
        double **A = calloc(nel, sizeof(double *)), *Aaux = calloc(144 * (size_t)nel, sizeof(double));
        for (I = 0; I < nel; I++) A[I] = &Aaux[144 * I];
        MD->A = A; MD->Aaux = Aaux;
        FF->assembly = ebe_assembly;
    } else {
        csr_Initialization(P, MD, FS);
        MD->AA = calloc(P->nnzero + 1, sizeof(double));
        MD->AA_bef = calloc(P->nnzero_bef + 1, sizeof(double));
        MD->AA_aft = calloc(P->nnzero_aft + 1, sizeof(double));
        rec("IA", 0, neq + 1, MD->IA);
        rec("JA", 0, P->nnzero, MD->JA);
        rec("perm", 0, neq, MD->Perm);
        int *sch = calloc(144 * (size_t)nel, sizeof(int));
        for (I = 0; I < nel; I++)
            for (J = 0; J < 144; J++) sch[144 * I + J] = MD->Scheme_by_Element[I][J];
        rec("scheme", 0, 144L * nel, sch);
        free(sch);
        FF->assembly = csr_assembly;
    }
    int *invp = calloc(4 * (size_t)nn, sizeof(int));
    for (I = 0; I < nn; I++)
        for (d = 0; d < 4; d++) {
            invp[4 * I + d] = Node[I].invP_id[d];
            if (Node[I].Type[d] != 0) {
                u[Node[I].invP_id[d]] = state[4 * I + d];
                du[Node[I].invP_id[d]] = incr[4 * I + d];
            }
        }
    rec("invP_id", 0, 4L * nn, invp);
    FF->gamma = h_gamma; FF->rhopresc = h_rho; FF->v1presc = h_v1; FF->v2presc = h_v2; FF->epresc = h_e;
    FF->Ax_Ay_calculations = dimless ? dimensionless_Ax_Ay_calculations : dimensional_Ax_Ay_calculations;
    FF->ShockCapture = shock ? Delta_CAU : h_noshock;
    Build_M_K_F_SUPG(P, MD, FS, FF);
    if (ebe) rec("A", 1, 144L * nel, MD->Aaux); else rec("AA", 1, P->nnzero + 1, MD->AA);
    rec("R", 1, neq + 1, R);
    if (ebe && hd[6] > 0) {
        /* the Euler linear solve of Predictor_Old.c:95-96 with the EBE scheme: BlockDiag_precond_EBE_setup
         * (nodal blocks, their inverses, R := D^-1 R), then pgmres with ebemvNDOF4 and BlockDiag_precond */
        BlockDiag_precond_EBE_setup(P, MD, FS, 1, R);
        rec("invBlockDiag", 1, 16L * nn, MD->invBlockDiag[0]);
        rec("MinvR", 1, neq, R);
        FF->mv = ebemvNDOF4;
        FF->precond = BlockDiag_precond;
        FF->precondR = NO_precond;
        P->KrylovBasisVectorsQuantity = hd[6];
        P->LinearMaxIter = hd[7];
        P->SolverTolerance = pr[5];
        P->iterations = 0;
        double *X = calloc(neq + 1, sizeof(double));
        pgmres(P, MD, FS, FF, R, X);
        int it = P->iterations;
        rec("iters", 0, 1, &it);
        rec("X", 1, neq, X);
        free(X);
    }
    fclose(g_out);
    MPI_Finalize();
    return 0;
}
