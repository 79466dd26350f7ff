/*
 * femgpu_compat.c -- the reference's hot-path entry points, by name and signature, forwarding to
 * libfemgpu.so. Compiled once per equation driver with that driver's macro and include path, like
 * the reference compiles 00_CommonFiles (MatrixVector_Operations/matvec.h:1-16), e.g.
 *
 *   gcc -c -DSSTranspEquation2D -I<reference>/00_CommonFiles -I<repo>/include -I/opt/rocm/include \
 *       -D__HIP_PLATFORM_AMD__ femgpu_compat.c
 *
 * and linked in place of the objects of MatrixVector_Operations/EBE/ebemv.c, .../CSR/csrmv.c,
 * Solvers_and_Preconditioners/pgmres.c and CSR/ILUp_precond(_setup).c (with -lfemgpu -lamdhip64).
 * The driver keeps its host arrays; the shim mirrors what the path needs on the device:
 *
 *   ILUp_precond_setup  uploads AA, factors on the device and applies F := M^-1 F (ILUp_precond_setup.c:8-61)
 *   ILUp_precond        z = U^-1 L^-1 p through the device sweeps (ILUp_precond.c:4-47)
 *   csrmv / ebemv       one device product (csrmv.c:4-65, ebemv.c:4-64)
 *   pgmres              the whole solve on the device: B up, X down, iterations added
 *                       (pgmres.c:5-200); matrix values are uploaded once per solve, never per step
 *
 *   SPARILU_clean / SPARMAT_clean  release the device objects (Postprocess.c:137-141 calls them
 *                       for the ILU preconditioners; the reference structures are never allocated)
 *
 * Navier-Stokes (9x9 blocks) and Euler (12x12) builds also provide the nodal block-diagonal
 * preconditioner (BlockDiagDOF3_precond(_EBE_setup) / BlockDiag_precond(_EBE_setup)) and, for Euler,
 * Block_scaling_EBE with the NO_scaling / NO_unscaling pair of the same file (Block_scaling.c); the
 * element matrices and F are uploaded, processed on the device and written back, so the driver's
 * host arrays stay the reference's state.
 *
 * Matrix values cross to the device when they change: every driver calls precond_setup after each
 * assembly (ST Process.c:60, Predictor_Old.c:95, NS Process.c:37), so the setup entry points --
 * ILUp_precond_setup, the block-diagonal setups and NO_precond_setup, which this shim also
 * provides -- mark the values stale, and the next product, preconditioner or solve uploads them
 * once. A product called before any setup uploads them on first use.
 *
 * Solver selection: the drivers pick pgmres by name (setSolver.c:6-15 accepts "GMRES" only), so the
 * mixed-precision solver is selected from the environment: FEMGPU_SOLVER=mixed makes pgmres run
 * femgpu_gmres_mixed (FP32 inner GMRES with kmax = KrylovBasisVectorsQuantity inside FP64
 * refinement, stopped at a true relative residual ||F - A x|| <= SolverTolerance ||F||) and makes
 * the setups leave F unpreconditioned (the mixed solver takes the raw right-hand side; the drivers
 * use F only as the solver's input). CSR systems use the parallel ILU(0) with Jacobi sweeps
 * (FEMGPU_MIXED_PREC=exact: the FP32 copy of the exact ILU(0)); EBE systems the FP32 copy of their
 * block-diagonal preconditioner.
 *
 * One rank only (the GPU holds the whole mesh): NPROC != 1 stops with a message. Failures print the
 * library's message and exit(1), as the reference does on fatal errors (mycalloc.c:10-13).
 * ILU(p) with p > 0 and preconditioners other than NO/ILU0/BlockDiag are not provided.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "MatrixVector_Operations/matvec.h"
#include "Solvers_and_Preconditioners/solvers.h"
#include "Solvers_and_Preconditioners/preconditioners.h"
#include "Solvers_and_Preconditioners/scaling.h"

#include "femgpu.h"

#if defined(SSNavierStokesEquations2D)
#define COMPAT_S 9
#elif defined(EulerEquations2D)
#define COMPAT_S 12
#else
#define COMPAT_S 3
#endif

static struct {
    femgpu_ctx *ctx;
    femgpu_csr *csr;
    femgpu_ebe *ebe;
    femgpu_blockdiag *bdiag;           /* nodal block diagonal (9x9 / 12x12 builds) */
    double *dA;                        /* element matrices on the device (block diagonal, scaling) */
    const MatrixDataType *csr_owner;   /* the MatrixData whose pattern csr holds */
    int factored;                      /* csr holds ILU(0) factors of the current AA */
    int stale;                         /* the driver's matrix values changed since the last upload */
    int mixed;                         /* FEMGPU_SOLVER=mixed */
    int mixed_exact;                   /* FEMGPU_MIXED_PREC=exact */
    double *dP, *dQ;                   /* NEQ + 1 doubles each */
    int NEQ;
} G;

static void fail(const char *where)
{
    printf("libfemgpu: %s: %s\n", where, femgpu_last_error());
    exit(1);
}

#define CALL(expr)                  \
    do {                            \
        if ((expr) != FEMGPU_OK)    \
            fail(#expr);            \
    } while (0)

static void compat_init(ParametersType *Parameters)
{
    if (G.ctx)
        return;
    if (Parameters->NPROC != 1) {
        printf("libfemgpu: one GPU holds the whole mesh; run with one MPI rank (NPROC = %d)\n",
               Parameters->NPROC);
        exit(1);
    }
    CALL(femgpu_create(0, NULL, &G.ctx));
    G.NEQ = Parameters->NEQ;
    G.stale = 1;
    const char *solver = getenv("FEMGPU_SOLVER"), *prec = getenv("FEMGPU_MIXED_PREC");
    G.mixed = solver && strcmp(solver, "mixed") == 0;
    G.mixed_exact = prec && strcmp(prec, "exact") == 0;
    if (hipMalloc((void **)&G.dP, sizeof(double) * (G.NEQ + 1)) != hipSuccess ||
        hipMalloc((void **)&G.dQ, sizeof(double) * (G.NEQ + 1)) != hipSuccess) {
        printf("libfemgpu: device allocation failed\n");
        exit(1);
    }
}

static void up(double *d, const double *h, int n)
{
    if (hipMemcpy(d, h, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        fail("upload");
}

static void down(double *h, const double *d, int n)
{
    if (hipMemcpy(h, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        fail("download");
}

/* The CSR operator of MatrixData (pattern uploaded on first use), its values uploaded when stale. */
static femgpu_csr *csr_of(ParametersType *Parameters, MatrixDataType *MatrixData)
{
    if (!G.csr || G.csr_owner != MatrixData) {
        if (G.csr)
            femgpu_csr_destroy(G.csr);
        CALL(femgpu_csr_create(G.ctx, Parameters->neq, Parameters->nnzero, MatrixData->IA, MatrixData->JA, &G.csr));
        G.csr_owner = MatrixData;
        G.factored = 0;
        G.stale = 1;
    }
    if (G.stale) {
        CALL(femgpu_csr_set_values(G.csr, MatrixData->AA, 0));
        G.factored = 0;
        G.stale = 0;
    }
    return G.csr;
}

/* The EBE operator (element lists on first use), its element matrices uploaded when stale. */
static femgpu_ebe *ebe_of(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs)
{
    if (!G.ebe) {
        CALL(femgpu_ebe_create(G.ctx, Parameters->nel, COMPAT_S, Parameters->NEQ, FemStructs->lmaux, &G.ebe));
        G.stale = 1;
    }
    if (G.stale) {
        CALL(femgpu_ebe_set_values(G.ebe, MatrixData->Aaux, 0));
        G.stale = 0;
    }
    return G.ebe;
}

/* NO_precond_setup (NO_precond_setup.c:3): nothing to set up; the matrix values the driver has just
 * assembled are marked for upload. */
int NO_precond_setup(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, int tag,
                     double *F)
{
    (void)MatrixData;
    (void)FemStructs;
    (void)tag;
    (void)F;
    compat_init(Parameters);
    G.stale = 1;
    return 0;
}

#if COMPAT_S > 3
#define COMPAT_ND (COMPAT_S / 3)

/* The node -> dof equations and node -> element lists, built on first use from the driver's Node
 * and Element arrays (prescribed dofs map to the sink NEQ). */
static femgpu_blockdiag *bdiag_of(ParametersType *Parameters, FemStructsType *FemStructs)
{
    if (G.bdiag)
        return G.bdiag;
    const int nn = Parameters->nnodes, nel = Parameters->nel;
    int *ien = malloc(sizeof(int) * 3 * (size_t)nel), *type = malloc(sizeof(int) * COMPAT_ND * (size_t)nn);
    int *invP = malloc(sizeof(int) * COMPAT_ND * (size_t)nn);
    if (!ien || !type || !invP) {
        printf("libfemgpu: host allocation failed\n");
        exit(1);
    }
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 3; a++)
            ien[3 * e + a] = FemStructs->Element[e].Vertex[a];
    for (int v = 0; v < nn; v++)
        for (int j = 0; j < COMPAT_ND; j++) {
            type[COMPAT_ND * v + j] = FemStructs->Node[v].Type[j] == 0 ? 0 : 1;
            invP[COMPAT_ND * v + j] = FemStructs->Node[v].Type[j] == 0 ? -1 : FemStructs->Node[v].invP_id[j];
        }
    CALL(femgpu_blockdiag_create(G.ctx, nn, nel, COMPAT_ND, Parameters->NEQ, ien, type, invP, &G.bdiag));
    free(ien);
    free(type);
    free(invP);
    return G.bdiag;
}

static double *elements_up(ParametersType *Parameters, MatrixDataType *MatrixData)
{
    const size_t n = (size_t)Parameters->nel * COMPAT_S * COMPAT_S;
    if (!G.dA && hipMalloc((void **)&G.dA, sizeof(double) * n) != hipSuccess)
        fail("element matrices: device allocation");
    if (hipMemcpy(G.dA, MatrixData->Aaux, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        fail("element matrices: upload");
    return G.dA;
}

/* BlockDiag(DOF3)_precond_EBE_setup: blocks of the current element matrices, F := D^-1 F (left raw
 * for the mixed solver). */
static int bdiag_setup(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *F)
{
    compat_init(Parameters);
    G.stale = 1;
    femgpu_blockdiag *B = bdiag_of(Parameters, FemStructs);
    CALL(femgpu_blockdiag_setup(B, elements_up(Parameters, MatrixData)));
    if (G.mixed)
        return 0;
    F[Parameters->NEQ] = 0.0;
    up(G.dP, F, Parameters->neq);
    CALL(femgpu_blockdiag_apply(B, G.dP, G.dP));
    down(F, G.dP, Parameters->neq);
    return 0;
}

/* BlockDiag(DOF3)_precond: z = D^-1 p (p[NEQ] and z[NEQ] set to 0, as the reference does). */
static int bdiag_apply(ParametersType *Parameters, FemStructsType *FemStructs, double *p, double *z)
{
    compat_init(Parameters);
    femgpu_blockdiag *B = bdiag_of(Parameters, FemStructs);
    p[Parameters->NEQ] = 0.0;
    up(G.dP, p, Parameters->neq);
    CALL(femgpu_blockdiag_apply(B, G.dP, G.dQ));
    down(z, G.dQ, Parameters->neq);
    z[Parameters->NEQ] = 0.0;
    return 0;
}

#if COMPAT_S == 12
int BlockDiag_precond_EBE_setup(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs,
                                int tag, double *F)
{
    (void)tag;
    return bdiag_setup(Parameters, MatrixData, FemStructs, F);
}

int BlockDiag_precond(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *p,
                      double *z)
{
    (void)MatrixData;
    return bdiag_apply(Parameters, FemStructs, p, z);
}

/* Block_scaling_EBE (Block_scaling.c:10-59): element matrices and F scaled on the device. */
int Block_scaling_EBE(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs)
{
    compat_init(Parameters);
    femgpu_blockdiag *B = bdiag_of(Parameters, FemStructs);
    double *dA = elements_up(Parameters, MatrixData);
    up(G.dP, FemStructs->F, Parameters->neq);
    CALL(femgpu_block_scaling(B, dA, G.dP));
    down(FemStructs->F, G.dP, Parameters->neq);
    FemStructs->F[Parameters->NEQ] = 0.0;
    if (hipMemcpy(MatrixData->Aaux, dA, sizeof(double) * (size_t)Parameters->nel * COMPAT_S * COMPAT_S,
                  hipMemcpyDeviceToHost) != hipSuccess)
        fail("element matrices: download");
    return 0;
}

int NO_scaling(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs)
{
    (void)Parameters;
    (void)MatrixData;
    (void)FemStructs;
    return 0;
}

int NO_unscaling(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *x)
{
    (void)Parameters;
    (void)MatrixData;
    (void)FemStructs;
    (void)x;
    return 0;
}
#else
int BlockDiagDOF3_precond_EBE_setup(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs,
                                    int tag, double *F)
{
    (void)tag;
    return bdiag_setup(Parameters, MatrixData, FemStructs, F);
}

int BlockDiagDOF3_precond(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs,
                          double *p, double *z)
{
    (void)MatrixData;
    return bdiag_apply(Parameters, FemStructs, p, z);
}
#endif
#endif

int csrmv(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *P, double *Q)
{
    (void)FemStructs;
    compat_init(Parameters);
    femgpu_csr *A = csr_of(Parameters, MatrixData);
    up(G.dP, P, Parameters->neq);
    CALL(femgpu_csrmv(A, G.dP, G.dQ));
    down(Q, G.dQ, Parameters->neq);
    return 0;
}

#if COMPAT_S == 3
int ebemv(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *P, double *Q)
#elif COMPAT_S == 9
int ebemvNDOF3(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *P, double *Q)
#else
int ebemvNDOF4(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *P, double *Q)
#endif
{
    compat_init(Parameters);
    femgpu_ebe *E = ebe_of(Parameters, MatrixData, FemStructs);
    up(G.dP, P, Parameters->NEQ);
    CALL(femgpu_ebemv(E, G.dP, G.dQ));
    down(Q, G.dQ, Parameters->NEQ);
    return 0;
}

int ILUp_precond_setup(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, int tag,
                       double *F)
{
    (void)FemStructs;
    (void)tag;   /* the pattern is uploaded once per MatrixData; values are refreshed on every call */
    compat_init(Parameters);
    if (atoi(Parameters->Preconditioner + 3) != 0) {
        printf("libfemgpu: only ILU(0) is provided (Preconditioner = %s)\n", Parameters->Preconditioner);
        exit(1);
    }
    G.stale = 1;
    femgpu_csr *A = csr_of(Parameters, MatrixData);
    if (G.mixed) {   /* the mixed solver factors in FP32 itself (or uses the exact factors) and takes F raw */
        if (G.mixed_exact)
            CALL(femgpu_ilu0_factor(A));
        G.factored = 1;
        return 0;
    }
    CALL(femgpu_ilu0_factor(A));
    G.factored = 1;
    up(G.dP, F, Parameters->neq);
    CALL(femgpu_ilu0_apply(A, G.dP, G.dP));
    down(F, G.dP, Parameters->neq);
    return 0;
}

int ILUp_precond(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs, double *p,
                 double *z)
{
    (void)FemStructs;
    compat_init(Parameters);
    femgpu_csr *A = csr_of(Parameters, MatrixData);
    if (!G.factored || G.mixed)
        fail("ILUp_precond before ILUp_precond_setup (or with FEMGPU_SOLVER=mixed, which keeps no FP64 factors)");
    up(G.dP, p, Parameters->neq);
    CALL(femgpu_ilu0_apply(A, G.dP, G.dQ));
    down(z, G.dQ, Parameters->neq);
    return 0;
}

int pgmres(ParametersType *Parameters, MatrixDataType *MatrixData, FemStructsType *FemStructs,
           FemFunctionsType *FemFunctions, double *B, double *X)
{
    femgpu_system S = {NULL, NULL, NULL, NULL};
    femgpu_solve_info info;
    compat_init(Parameters);
    if (FemFunctions->precond == ILUp_precond) {
        S.ilu = csr_of(Parameters, MatrixData);
        if (!G.factored)
            fail("pgmres: ILU(0) preconditioner used before ILUp_precond_setup");
    }
#if COMPAT_S == 12
    else if (FemFunctions->precond == BlockDiag_precond)
        S.bdiag = bdiag_of(Parameters, FemStructs);
#elif COMPAT_S == 9
    else if (FemFunctions->precond == BlockDiagDOF3_precond)
        S.bdiag = bdiag_of(Parameters, FemStructs);
#endif
    else if (FemFunctions->precond != NO_precond) {
        printf("libfemgpu: preconditioner %s is not provided\n", Parameters->Preconditioner);
        exit(1);
    }
    if (FemFunctions->mv == csrmv)
        S.csr = csr_of(Parameters, MatrixData);
#if COMPAT_S == 3
    else if (FemFunctions->mv == ebemv)
#elif COMPAT_S == 9
    else if (FemFunctions->mv == ebemvNDOF3)
#else
    else if (FemFunctions->mv == ebemvNDOF4)
#endif
        S.ebe = ebe_of(Parameters, MatrixData, FemStructs);
    else {
        printf("libfemgpu: matrix-vector product %s is not provided; use EBE or CSR\n",
               Parameters->MatrixVectorProductScheme);
        exit(1);
    }
    up(G.dP, B, Parameters->neq);
    if (G.mixed) {
        femgpu_mixed_opts mo;
        memset(&mo, 0, sizeof(mo));
        mo.inner_kmax = Parameters->KrylovBasisVectorsQuantity;
        mo.inner_cycles = 1;
        mo.inner_tol = 1e-4;
        mo.max_refinements = 200;
        if (S.csr && !S.ilu) {
            printf("libfemgpu: FEMGPU_SOLVER=mixed needs a preconditioner (ILU0 for CSR, BlockDiag for EBE)\n");
            exit(1);
        }
        if (S.ilu) {
            mo.sweep_mode = G.mixed_exact ? 0 : 2;
            mo.factor_sweeps = G.mixed_exact ? 0 : 2;
        }
        CALL(femgpu_gmres_mixed(G.ctx, &S, &mo, Parameters->SolverTolerance, G.dP, G.dQ, &info));
    } else {
        CALL(femgpu_pgmres(G.ctx, &S, Parameters->KrylovBasisVectorsQuantity, Parameters->LinearMaxIter,
                           Parameters->SolverTolerance, G.dP, G.dQ, &info));
    }
    down(X, G.dQ, Parameters->neq);
    Parameters->iterations += info.iterations;
    return 0;
}

/* The ILU(0) data lives on the device: Postprocess's clean-up of the reference's ILU structures
 * (which this shim never allocates) releases the library objects instead. */
void SPARILU_clean(SparILU *lu)
{
    (void)lu;
    if (G.csr)
        femgpu_csr_destroy(G.csr);
    if (G.ebe)
        femgpu_ebe_destroy(G.ebe);
    if (G.bdiag)
        femgpu_blockdiag_destroy(G.bdiag);
    G.csr = NULL;
    G.ebe = NULL;
    G.bdiag = NULL;
    G.csr_owner = NULL;
    G.factored = 0;
}

void SPARMAT_clean(SparMAT *mat)
{
    (void)mat;
}
