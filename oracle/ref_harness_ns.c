/*
 * ref_harness_ns.c -- one-rank driver around the UNMODIFIED reference Navier-Stokes assembly
 * sources of 04_SS_NavierStokes_Equations (compiled in place from the reference tree by
 * oracle/ref.mk). Test infrastructure: it writes golden vectors for tests/golden; it is never
 * shipped or run on a GPU machine.
 *
 * It plays the role of Preprocess.c for NPROC = 1 and calls, once, the reference's
 * Build_K_F_SUPG_PSPG (with eval_U, setzeros, Fill_LM, csr_Initialization and ebe/csr_assembly)
 * on a given state. It supplies the problem callbacks (v1presc, v2presc, ppresc, f1ext, f2ext) as
 * per-node tables looked up by exact coordinates, and MPI_VectorUpdate for one rank.
 *
 * Input (binary, little endian): int32 header[16] {nn, nel, ebe, reorder}, float64 params[16]
 * {Re}, then x[nn], y[nn], type[3nn] (int32), ien[3nel] (int32), presc[3nn], fsrc[2nn],
 * state[3nn] (float64; the value of each unknown dof).
 * Output: records {char name[32]; int32 kind (0 int32, 1 float64); int64 count; payload}.
 */
#include "SSNavierStokesEquations.h"

int Fill_LM(ParametersType *, int **, NodeType *, ElementType *);
void csr_Initialization(ParametersType *, MatrixDataType *, FemStructsType *);
void ebe_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[9]);
void csr_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[9]);
int Build_K_F_SUPG_PSPG(ParametersType *, MatrixDataType *, FemStructsType *, FemFunctionsType *);

static int g_nn;
static const double *g_x, *g_y, *g_presc, *g_fsrc;

static int node_at(double X, double Y)
{
    for (int i = 0; i < g_nn; i++)
        if (g_x[i] == X && g_y[i] == Y) return i;
    return -1;
}
static double h_v1(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[3 * i]; }
static double h_v2(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[3 * i + 1]; }
static double h_p(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_presc[3 * i + 2]; }
static double h_f1(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_fsrc[2 * i]; }
static double h_f2(double X, double Y) { int i = node_at(X, Y); return i < 0 ? 0.0 : g_fsrc[2 * i + 1]; }

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
    const int nn = hd[0], nel = hd[1], ebe = hd[2], reorder = hd[3];
    double *x = calloc(nn, sizeof(double)), *y = calloc(nn, sizeof(double));
    double *presc = calloc(3 * (size_t)nn, sizeof(double)), *fsrc = calloc(2 * (size_t)nn, sizeof(double));
    double *state = calloc(3 * (size_t)nn, sizeof(double));
    int *type = calloc(3 * (size_t)nn, sizeof(int)), *ien = calloc(3 * (size_t)nel, sizeof(int));
    size_t ok = fread(x, sizeof(double), nn, in) + fread(y, sizeof(double), nn, in) +
                fread(type, sizeof(int), 3 * (size_t)nn, in) + fread(ien, sizeof(int), 3 * (size_t)nel, in) +
                fread(presc, sizeof(double), 3 * (size_t)nn, in) + fread(fsrc, sizeof(double), 2 * (size_t)nn, in) +
                fread(state, sizeof(double), 3 * (size_t)nn, in);
    fclose(in);
    if (ok != (size_t)13 * nn + 3 * (size_t)nel) MPI_Abort(MPI_COMM_WORLD, 4);
    g_nn = nn; g_x = x; g_y = y; g_presc = presc; g_fsrc = fsrc;
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
        for (d = 0; d < 3; d++) {
            Node[I].Type[d] = type[3 * I + d];
            Node[I].Send[d] = 1;
            Node[I].invP_id[d] = (type[3 * I + d] == 0) ? -1 : neq++;
        }
    }
    for (I = 0; I < nel; I++)
        for (J = 0; J < 3; J++) Element[I].Vertex[J] = ien[3 * I + J];
    P->neq = neq; P->neq_bef = 0; P->NEQ = neq; P->nel = nel; P->nnodes = nn; P->RANK = 0; P->NPROC = 1;
    P->ReynoldsNumber = pr[0];
    strcpy(P->reordering, reorder ? "SYMRCM" : "NOT");
    strcpy(P->MatrixVectorProductScheme, ebe ? "EBE" : "CSR");
    double *F = calloc(neq + 1, sizeof(double)), *u = calloc(neq + 1, sizeof(double));
    int **lm = calloc(nel, sizeof(int *)), *lmaux = calloc(9 * (size_t)nel, sizeof(int));
    for (I = 0; I < nel; I++) lm[I] = &lmaux[9 * I];
    Fill_LM(P, lm, Node, Element);
    FS->Node = Node; FS->Element = Element; FS->lm = lm; FS->lmaux = lmaux;
    FS->F = F; FS->u = u;
    rec("lm", 0, 9L * nel, lmaux);
    if (ebe) {
        double **A = calloc(nel, sizeof(double *)), *Aaux = calloc(81 * (size_t)nel, sizeof(double));
        for (I = 0; I < nel; I++) A[I] = &Aaux[81 * I];
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
        int *sch = calloc(81 * (size_t)nel, sizeof(int));
        for (I = 0; I < nel; I++)
            for (J = 0; J < 81; J++) sch[81 * I + J] = MD->Scheme_by_Element[I][J];
        rec("scheme", 0, 81L * nel, sch);
        free(sch);
        FF->assembly = csr_assembly;
    }
    /* the state: each unknown dof takes its node's state value (after any renumbering) */
    int *invp = calloc(3 * (size_t)nn, sizeof(int));
    for (I = 0; I < nn; I++)
        for (d = 0; d < 3; d++) {
            invp[3 * I + d] = Node[I].invP_id[d];
            if (Node[I].Type[d] != 0) u[Node[I].invP_id[d]] = state[3 * I + d];
        }
    rec("invP_id", 0, 3L * nn, invp);
    FF->v1presc = h_v1; FF->v2presc = h_v2; FF->ppresc = h_p; FF->f1ext = h_f1; FF->f2ext = h_f2;
    Build_K_F_SUPG_PSPG(P, MD, FS, FF);
    if (ebe) rec("A", 1, 81L * nel, MD->Aaux); else rec("AA", 1, P->nnzero + 1, MD->AA);
    rec("F", 1, neq + 1, F);
    fclose(g_out);
    MPI_Finalize();
    return 0;
}
