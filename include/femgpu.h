/*
 * femgpu.h -- C ABI of libfemgpu.so, the MI355X (gfx950) implementation of the linear-solve hot
 * path of PARALLEL_FEM_CODES (steady/transient SUPG transport, NDOF = 1, plus block EBE).
 *
 * Plain C types only: ints, doubles, floats, raw host or device pointers and opaque handles.
 * Every function returns FEMGPU_OK (0) or a negative status; femgpu_last_error() describes the
 * last failure of the calling thread. Device pointers ("d_" prefix) are HIP device memory on the
 * context's device; host pointers are ordinary memory. Indices are 0-based, one rank
 * (neq_bef = 0, NEQ = neq): vectors have neq entries plus, where the reference uses it, the
 * Dirichlet sink at index NEQ.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   ebemv / ebemvNDOF3 / ebemvNDOF4 / csrmv  00_CommonFiles/MatrixVector_Operations/matvec.h:3-18
 *   pgmres                                   00_CommonFiles/Solvers_and_Preconditioners/solvers.h:14
 *   ILUp_precond / ILUp_precond_setup        00_CommonFiles/Solvers_and_Preconditioners/preconditioners.h:30,35
 *   ddot / daxpy / dcopy / dscal / dzero     00_CommonFiles/BLAS_Operations/ourBLAS.h:1-5
 *   Fill_LM / csr_Initialization / Permutation_of_LM
 *                                            01_SS_Transport_Equation/2D/01_CommonFiles/SSTranspEquation.h:174,176,225
 *   Build_K_F_SUPG / Build_M_K_R_SUPG        SSTranspEquation.h:217, 02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c:6
 * The reference-signature shims that forward to this ABI live in parallel_fem_codes_amd/compat/.
 */
#ifndef FEMGPU_H
#define FEMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define FEMGPU_OK 0
#define FEMGPU_ERR_ARG -1         /* bad argument (size, null pointer, unknown option) */
#define FEMGPU_ERR_HIP -2         /* HIP runtime failure */
#define FEMGPU_ERR_ZERO_PIVOT -3  /* ILU(0) met a zero pivot (the reference exits, ILUp_precond_setup.c:155-164) */
#define FEMGPU_ERR_NOMEM -4       /* device allocation failed */
#define FEMGPU_ERR_STATE -5       /* object used before it was set up (e.g. apply before factor) */
#define FEMGPU_ERR_TIMEOUT -6     /* a bounded device-side wait expired (sweep scheduler) */

const char *femgpu_last_error(void);
const char *femgpu_version(void);

/* =========================================================== integer structure work (host) */

/* Unknown numbering of Parallel_Preprocess with reordering NOT and one part
 * (Parallel_Preprocess/Preprocess.c:37-53): running count of Type-1 dofs, -1 for Dirichlet.
 * type, invP_id: nnodes*ndof node-major. Returns neq (>= 0) or a negative status. */
int femgpu_number_unknowns(int nnodes, int ndof, const int *type, int *invP_id);

/* Fill_LM (01_SS_Transport_Equation/2D/01_CommonFiles/Fill_LM.c:3-27;
 * 04_SS_NavierStokes_Equations/2D/01_CommonFiles/Fill_LM.c:3-31 for ndof > 1). */
int femgpu_fill_lm(int nel, int ndof, int NEQ, const int *ien, const int *type,
                   const int *invP_id, int *lm);

/* Nonzero count of the CSR block csr_Initialization builds from lm (csr_Initialization.c:36-99).
 * s = 3*ndof. Returns nnz or a negative status. */
int femgpu_csr_nnz(int nel, int s, int neq, const int *lm);

/* csr_Initialization for one rank (01_SS_Transport_Equation/2D/01_CommonFiles/csr_Initialization.c:5-571):
 * IA[neq+1], JA[nnz], scheme[nel*s*s] (slot of each element entry, nnz = dummy), perm[neq]
 * (new -> old). reorder: 0 = NOT, 1 = SYMRCM, 2 = Spectral (femgpu_spectral).
 * invP_id of Type-1 dofs is renumbered; lm too when permute_lm != 0 (the transient driver's
 * variant, 02_Transport_Equation/2D/01_CommonFiles/csr_Initialization.c:452-454). */
int femgpu_csr_init(int nel, int s, int neq, int *lm, int reorder, int permute_lm,
                    int ndofs_total, const int *type, int *invP_id, int *IA, int *JA, int *scheme,
                    int *perm);

/* REORDERING_SYMRCM (00_CommonFiles/Reordering/symrcm.c:4-237). perm[new] = old. */
int femgpu_symrcm(int n, const int *IA, const int *JA, int *perm);

/* REORDERING_SPECTRAL (00_CommonFiles/Reordering/reordering.c:27-77): perm[new] = old, the
 * unknowns by ascending entry of the Fiedler vector of the pattern's graph Laplacian (ties by
 * index). The reference takes the vector from HSL_MC73 (third party); femgpu_fiedler computes it
 * independently (multilevel: heavy-edge coarsening, dense Jacobi eigensolve of the coarsest
 * graph, power-iteration refinement with refine_iters steps per level; femgpu_spectral uses 60),
 * so the permutation is not the reference's bit for bit. fvec: n entries, zero mean, unit norm. */
int femgpu_fiedler(int n, const int *IA, const int *JA, int refine_iters, double *fvec);
int femgpu_spectral(int n, const int *IA, const int *JA, int *perm);

/* Permutation_of_LM (01_SS_Transport_Equation/2D/01_CommonFiles/Permutation_of_LM.c:37-105):
 * Perm/invPerm of NEQ+1 entries, Perm[NEQ] = invPerm[NEQ] = NEQ. */
int femgpu_ebe_permutation(int nel, int NEQ, const int *ien, const int *type, const int *invP_id,
                           int reorder, int *Perm, int *invPerm);

/* Greedy element colouring (no reference counterpart: it replaces the serial scatter of
 * ebemv.c:40-55): elements in index order take the smallest colour not used by an element that
 * shares an unknown (lm entry < NEQ). Returns the number of colours, or a negative status. */
int femgpu_ebe_colour(int nel, int s, int NEQ, const int *lm, int *colour);

/* Dependency levels of the ILU(0) sweeps (no reference counterpart: the reference sweeps
 * serially, ILUp_precond.c:26-44): lower != 0: level(i) = 1 + max level(j), j < i in row i,
 * else 0; lower == 0: the same over j > i from the last row. Returns depth (max level + 1). */
int femgpu_ilu_levels(int n, const int *IA, const int *JA, int lower, int *level);

/* =========================================================== device context */

typedef struct femgpu_ctx femgpu_ctx;

/* stream: a hipStream_t (cast to void*) or NULL for the device's null stream. */
int femgpu_create(int device, void *stream, femgpu_ctx **out);
int femgpu_destroy(femgpu_ctx *ctx);
int femgpu_sync(femgpu_ctx *ctx);

/* =========================================================== CSR operator + ILU(0) */

typedef struct femgpu_csr femgpu_csr;

/* Upload a CSR structure (host IA/JA, rows sorted) and prepare the level-scheduled ILU(0)
 * sweeps. Values are set separately. */
int femgpu_csr_create(femgpu_ctx *ctx, int n, int nnz, const int *IA, const int *JA,
                      femgpu_csr **out);
/* As femgpu_csr_create with a choice of the plans built: FEMGPU_PLANS_FP64 (exact ILU(0)
 * factorisation and level-scheduled sweeps, needed by femgpu_ilu0_factor / _apply, pgmres with
 * ILU(0) and the mixed solver with sweep_mode 0) and FEMGPU_PLANS_MIXED (parallel FP32 ILU(0) and
 * Jacobi iterations). femgpu_csr_create builds both; the serving loop builds only the mixed ones.
 * No reference counterpart. */
#define FEMGPU_PLANS_FP64 1
#define FEMGPU_PLANS_MIXED 2
int femgpu_csr_create_ex(femgpu_ctx *ctx, int n, int nnz, const int *IA, const int *JA, int plans,
                         femgpu_csr **out);
/* Re-plans A for a new structure (same plan flags), reusing its device buffers where they are large
 * enough: for callers that solve a stream of different systems (the request loop) without a
 * hipMalloc/hipFree per system. Values, factors and captured graphs of the old structure are void.
 * No reference counterpart. */
int femgpu_csr_rebuild(femgpu_csr *A, int n, int nnz, const int *IA, const int *JA);
int femgpu_csr_destroy(femgpu_csr *A);
/* AA: nnz values (the reference's dummy slot AA[nnz] is not needed); on_device selects whether
 * AA is a host or a device pointer. Also refreshes the FP32 copy. */
int femgpu_csr_set_values(femgpu_csr *A, const double *AA, int on_device);
/* Device pointer of the FP64 value array (nnz+1 entries, last = dummy slot), for in-place
 * assembly; call femgpu_csr_values_changed afterwards to refresh the FP32 copy. */
int femgpu_csr_values_ptr(femgpu_csr *A, double **d_AA);
int femgpu_csr_values_changed(femgpu_csr *A);
/* ILU(0) sweep scheduler description (no reference counterpart: the reference sweeps serially). */
typedef struct femgpu_sweep_info {
    int kernel;            /* 0: streamed single-workgroup sweep, 1: packed multi-workgroup sweep */
    int depth_lower;       /* dependency levels of the L sweep */
    int depth_upper;       /* dependency levels of the U sweep */
    int steps_lower;       /* streamed: barrier steps of the L sweep (levels split to <= 512 rows) */
    int steps_upper;
    int ring;              /* streamed: LDS ring entries holding recent results */
    int entries_per_row;   /* streamed: padded dependency entries per row */
    int chunks_lower;      /* packed: row chunks of the L sweep */
    int rows_per_chunk;    /* packed: row bound per chunk */
    int lds_limit;         /* LDS bytes available to one workgroup */
    int factor_kernel;     /* 0: streamed single-workgroup factorisation, 1: level-synchronous lists */
    int jacobi_column_bytes;   /* Jacobi sliced-ELL triangles: 2 (16-bit distances from the row), 4, 0: none */
    int sell_column_bytes;     /* sliced-ELL products: 2 (16-bit distances from the row), 4, 0: none */
    int long_slots;            /* streamed: LDS slots for results read beyond the ring (larger of L, U) */
    int factor_compact;        /* streamed factorisation: 1 for the compact row shape (every row at most
                                  7 entries, 3 L and 3 U), 2 for the compact shape on the pair pattern
                                  of the structured triangle meshes, 0 for the general one */
} femgpu_sweep_info;
int femgpu_csr_sweep_info(femgpu_csr *A, femgpu_sweep_info *info);
/* Select the sweep kernel: 0 = streamed when the plan allows it (default), 1 = packed. */
int femgpu_csr_set_sweep_kernel(femgpu_csr *A, int kernel);
/* Select the numeric factorisation kernel: 0 = streamed when the plan allows it (default),
 * 1 = level-synchronous update lists. Both give the same factors. */
int femgpu_csr_set_factor_kernel(femgpu_csr *A, int kernel);

/* csrmv (CSR/csrmv.c:4-65), local block: Q = A P. FP64 and FP32 (FP32 copy of the values). */
int femgpu_csrmv(femgpu_csr *A, const double *d_P, double *d_Q);
int femgpu_csrmv_f32(femgpu_csr *A, const float *d_P, float *d_Q);

/* Kernel of the FP64 products (csrmv, pgmres's operator, the mixed solver's FP64 residual):
 * L = 1, 2, 4, 8, 16 lanes per row of the CSR kernel (64 / L rows per wavefront), or L = 0: one row
 * per thread on the 64-row sliced-ELL copy, with an FP64 copy of the values in that layout (kept,
 * and refreshed by femgpu_csr_values_changed, only while L = 0; needs the sliced-ELL copy with
 * slices at most 16 wide, else FEMGPU_ERR_ARG). femgpu_csr_autotune times every candidate on A's
 * structure and current values (20 launches each, HIP events) and keeps the fastest; us[] =
 * microseconds per launch for L = 1, 2, 4, 8, 16, us_sell for L = 0 (0 where not possible).
 * Synchronises the context's stream. Default L = 8. No reference counterpart (the reference's
 * csrmv is one loop per row). */
typedef struct {
    int lanes;             /* the L chosen */
    int pad;
    double us[5];          /* per-launch time of L = 1, 2, 4, 8, 16 */
    double us_sell;        /* per-launch time of L = 0 */
} femgpu_tune_info;
int femgpu_csr_autotune(femgpu_csr *A, femgpu_tune_info *info);
int femgpu_csr_set_lanes(femgpu_csr *A, int lanes);

/* ILU(0) numeric factorisation of the current values (ILUp_precond_setup.c:67-168 with p = 0),
 * level-scheduled on the device; also builds the FP32 factors. */
int femgpu_ilu0_factor(femgpu_csr *A);
/* Factor on the host from the current values and upload (same arithmetic, serial order). */
int femgpu_ilu0_factor_host(femgpu_csr *A);
/* ILUp_precond (CSR/ILUp_precond.c:4-47): z = U^-1 L^-1 p; d_p == d_z allowed. With an apply mode
 * set by femgpu_ilu0_set_apply, the FP64 Jacobi-applied parallel factors instead. */
int femgpu_ilu0_apply(femgpu_csr *A, const double *d_p, double *d_z);
/* FP64 ILU(0) factors by `sweeps` fixed-point sweeps of the parallel ILU(0) iteration (Chow and
 * Patel, SIAM J. Sci. Comput. 37(2), 2015) from the current values; no exact factorisation. Needs
 * rows and strict triangles of at most 16 entries; 1 <= sweeps <= 1024. No reference counterpart:
 * the FP64 form of the mixed solver's preconditioner (a precision-only comparison with it). */
int femgpu_ilu0_factor_parallel(femgpu_csr *A, int sweeps);
/* How the FP64 ILU(0) preconditioner of A is applied (femgpu_ilu0_apply, and pgmres / pcg with A as
 * preconditioner): 0 = the exact level-scheduled sweeps on the factors of femgpu_ilu0_factor (the
 * default, ILUp_precond); k > 0 = k Jacobi iterations per triangle on the factors of
 * femgpu_ilu0_factor_parallel. No reference counterpart. */
int femgpu_ilu0_set_apply(femgpu_csr *A, int jacobi_iterations);
int femgpu_ilu0_apply_f32(femgpu_csr *A, const float *d_p, float *d_z);
/* Copy the factors to the host: LU[nnz] in A's layout (L multipliers below, U above the
 * diagonal), Dinv[n] = 1/pivot. */
int femgpu_ilu0_get(femgpu_csr *A, double *LU, double *Dinv);

/* FP32 preconditioner pieces of the mixed solver (no reference counterpart).
 * Parallel ILU(0): FP32 factors in A's layout by `sweeps` fixed-point sweeps of the Chow-Patel
 * iteration from the current values (converges to the ILU(0) factors; at the fixed point after
 * twice as many sweeps as the factorisation has dependency levels). */
int femgpu_ilu0_factor_parallel_f32(femgpu_csr *A, int sweeps);
/* FP32 copy of the factors of femgpu_ilu0_factor into the same FP32 factor storage. */
int femgpu_ilu0_factor_to_f32(femgpu_csr *A);
/* z ~ U^-1 L^-1 p with `iterations` (1..1024) Jacobi iterations per triangle on the FP32 factors
 * (exact after as many iterations as the sweep has dependency levels); d_p == d_z allowed. */
int femgpu_ilu0_apply_jacobi_f32(femgpu_csr *A, int iterations, const float *d_p, float *d_z);
/* One Jacobi iteration on one triangle of the FP32 factors (the kernels femgpu_ilu0_apply_jacobi_f32
 * chains): upper = 0: d_out = d_rhs - L_s d_in; upper = 1: d_out = Dinv (d_rhs - U_s d_in). d_in
 * NULL stands for the starting iterate (d_rhs for L, Dinv d_rhs for U); d_out must not alias d_in
 * or d_rhs. No reference counterpart (the preconditioner's building block, for callers that
 * schedule the iterations themselves). */
int femgpu_ilu0_jacobi_iteration_f32(femgpu_csr *A, int upper, const float *d_rhs, const float *d_in,
                                     float *d_out);
/* Copy the FP32 factors to the host (layout as femgpu_ilu0_get). */
int femgpu_ilu0_get_f32(femgpu_csr *A, float *LU, float *Dinv);

/* =========================================================== EBE operator */

typedef struct femgpu_ebe femgpu_ebe;

/* Element-by-element operator with s x s blocks (s = 3 transport, 9 Navier-Stokes, 12 Euler),
 * lm[nel*s] host array with Dirichlet sink NEQ, lm[e][ndof*a + d] = unknown of dof d at vertex a
 * (ndof = s/3; Fill_LM's layout). The product is gathered per node: each node's element rows are
 * stored together in element order, so the sums keep the reference's order with no atomics. */
int femgpu_ebe_create(femgpu_ctx *ctx, int nel, int s, int NEQ, const int *lm, femgpu_ebe **out);
int femgpu_ebe_destroy(femgpu_ebe *E);
/* A: nel*s*s row-major element matrices (ebe_assembly.c layout); host or device pointer. Fills the
 * FP64 operator and its FP32 copy (the mixed-precision inner solver's). */
int femgpu_ebe_set_values(femgpu_ebe *E, const double *A, int on_device);

typedef struct {
    int groups;          /* nodes with at least one unknown (one thread of the product each) */
    int max_incidences;  /* most elements at one node */
    int column_bytes;    /* 2: int16 offsets from the node's first unknown; 4: int32 unknowns */
    int pad;
    long long bytes_f64; /* bytes one FP64 product moves in this layout (values, columns, P, Q) */
    long long bytes_f32; /* the same for the FP32 product */
} femgpu_ebe_info;
int femgpu_ebe_get_info(femgpu_ebe *E, femgpu_ebe_info *info);
/* ebemv / ebemvNDOF3 / ebemvNDOF4 (MatrixVector_Operations/EBE/ebemv.c:4-64, ebemvNDOF3.c:3-92,
 * ebemvNDOF4.c:4-121): Q[0..NEQ) = sum_e A_e P|_e with P[NEQ] read as 0, bit-identical to the
 * reference's element loop. */
int femgpu_ebemv(femgpu_ebe *E, const double *d_P, double *d_Q);
/* The same product on the FP32 copy of the element matrices (no reference counterpart: the
 * mixed-precision inner solver's operator). */
int femgpu_ebemv_f32(femgpu_ebe *E, const float *d_P, float *d_Q);

/* =========================================================== BLAS on device vectors */

int femgpu_ddot(femgpu_ctx *ctx, int n, const double *d_x, const double *d_y, double *result);
/* No reference counterpart: for callers that cannot allocate (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc). femgpu_ddot's reduction (the same kernels, the
 * same result bits) with the result written to device memory *d_result on the context's stream,
 * no host synchronisation, and its partial sums in a fixed 1024-entry (8 KiB) array per GPU inside
 * the library instead of an allocation: it calls no allocator. Any n >= 0 fits (the reduction uses
 * at most 1024 blocks). Calls that overlap on one GPU share that array and race. */
int femgpu_ddot_noalloc(femgpu_ctx *ctx, int n, const double *d_x, const double *d_y, double *d_result);
int femgpu_daxpy(femgpu_ctx *ctx, int n, double a, const double *d_x, double *d_y);
int femgpu_dscal(femgpu_ctx *ctx, int n, double a, double *d_x);
int femgpu_dcopy(femgpu_ctx *ctx, int n, const double *d_x, double *d_y);
int femgpu_dzero(femgpu_ctx *ctx, int n, double *d_x);

/* =========================================================== solvers */

typedef struct femgpu_blockdiag femgpu_blockdiag;

typedef struct {
    femgpu_csr *csr;   /* operator, when non-null */
    femgpu_ebe *ebe;   /* operator, when csr is null */
    femgpu_csr *ilu;   /* factored ILU(0) preconditioner (left), or NULL for none */
    femgpu_blockdiag *bdiag;   /* nodal block-diagonal preconditioner (left), used when ilu is NULL */
} femgpu_system;

typedef struct {
    int iterations;    /* Arnoldi steps (Parameters->iterations increment in pgmres.c:184) */
    int cycles;        /* restart cycles */
    int inner_iterations; /* mixed precision: total FP32 Arnoldi steps */
    int status;        /* 0 converged, 1 iteration cap reached */
    double rho;        /* final Givens estimate of the (preconditioned) residual norm */
    double normb;      /* norm of the right-hand side the stopping test refers to */
    double true_rel_residual;  /* ||F - A x|| / ||F|| in FP64 (mixed solver; 0 otherwise) */
    double prec_rel_residual;  /* ||M^-1(F - A x)|| / ||M^-1 F|| in FP64 (mixed solver) */
} femgpu_solve_info;

/* pgmres (00_CommonFiles/Solvers_and_Preconditioners/pgmres.c:5-200): restarted, left
 * preconditioned GMRES with modified Gram-Schmidt and Givens rotations; kmax-1 Arnoldi steps per
 * cycle, at most lmax cycles, X zeroed on entry, stop when |e_{i+1}| <= tol*||B||. B is the
 * already preconditioned right-hand side (as precond_setup leaves it). FP64 throughout. */
int femgpu_pgmres(femgpu_ctx *ctx, const femgpu_system *sys, int kmax, int lmax, double tol,
                  const double *d_B, double *d_X, femgpu_solve_info *info);

typedef struct {
    int inner_kmax;        /* basis size of the FP32 inner GMRES (kmax-1 steps per cycle) */
    int inner_cycles;      /* inner restart cycles per refinement step */
    double inner_tol;      /* relative tolerance of each inner solve (preconditioned norm) */
    int max_refinements;   /* outer FP64 refinement steps */
    int sweep_mode;        /* FP32 ILU application: 0 exact level-scheduled sweeps,
                              k > 0: k Jacobi iterations per triangular factor */
    int factor_sweeps;     /* FP32 factors: 0 = the FP32 copy of the factorisation done by
                              femgpu_ilu0_factor; k > 0 = k sweeps of the parallel (fixed-point)
                              ILU(0) from the current values, no FP64 factorisation needed
                              (requires sweep_mode > 0) */
    int orthogonalization; /* inner Arnoldi: 0 modified Gram-Schmidt (one pass per basis vector),
                              1 classical Gram-Schmidt (all coefficients from one pass, one
                              update pass: fewer passes over the basis) */
    int graphs;            /* 1: every Arnoldi step of the inner solver (SpMV, preconditioner,
                              orthogonalization) is captured once into a hipGraph and replayed;
                              the graphs are kept by the context and re-captured when the matrix
                              or the workspace changes. Same kernels, same results. */
    int pad;
} femgpu_mixed_opts;

/* Mixed precision: FP64 iterative refinement (residual F - A x, update x += d, Givens and least
 * squares in FP64) around an FP32 left-preconditioned inner GMRES. Stops when
 * ||F - A x|| <= tol * ||F|| (FP64-verified). F is the raw (unpreconditioned) right-hand side.
 * Two system forms: sys->csr with sys->ilu (ILU(0), factored unless opt->factor_sweeps > 0; the
 * inner solver uses the FP32 sliced-ELL copy and FP32 factors), or sys->ebe with sys->bdiag (set up
 * on the current element matrices; the inner solver uses the FP32 element matrices and FP32 inverse
 * blocks, sweep_mode / factor_sweeps ignored, graphs must be 0). X = 0 is returned where it already
 * passes (||F|| = 0 or tol >= 1: status 0), with max_refinements = 0 (status 1 unless ||F|| = 0)
 * and for a non-finite F (status 1). No reference counterpart. */
int femgpu_gmres_mixed(femgpu_ctx *ctx, const femgpu_system *sys, const femgpu_mixed_opts *opt,
                       double tol, const double *d_F, double *d_X, femgpu_solve_info *info);

typedef struct {
    int refinements;       /* outer FP64 refinement steps while the system was active */
    int inner_iterations;  /* FP32 Arnoldi steps of the cycles the system took part in */
    int status;            /* 0 converged, 1 refinement cap reached */
    int pad;
    double normF;          /* ||F_s|| */
    double true_rel_residual;  /* ||F_s - A_s x_s|| / ||F_s|| in FP64 */
} femgpu_batch_info;

/* Mixed precision for nsys independent systems stored as one block-diagonal CSR matrix (system s on
 * rows [row_off[s], row_off[s+1]), host array, row_off[0] = 0, row_off[nsys] = n; no entries
 * across blocks) with its ILU(0) preconditioner (block diagonal as well). FP64 refinement per
 * system: each system's residual is verified in FP64 and the system is masked out (its block of
 * the inner right-hand side zeroed, so every Krylov vector is exactly zero there) as soon as
 * ||F_s - A_s x_s|| <= tol ||F_s||; active systems enter each inner FP32 GMRES cycle with their
 * residual scaled to unit norm, and the cycle's correction is scaled back per system. One launch
 * sequence serves the whole batch. info: nsys entries. No reference counterpart. */
int femgpu_gmres_mixed_batched(femgpu_ctx *ctx, const femgpu_system *sys, int nsys, const int *row_off,
                               const femgpu_mixed_opts *opt, double tol, const double *d_F, double *d_X,
                               femgpu_batch_info *info);

/* The same batched mixed-precision solve as one persistent launch with a workgroup per system:
 * each system runs its own FP64 refinement and FP32 GMRES(inner_kmax) (its own Krylov space and
 * Hessenberg matrix, fused MGS, Givens and least squares in LDS) on the sliced-ELL copy of A with
 * the parallel ILU(0) factors applied by Jacobi iterations, so a system's answer is bit-identical
 * alone and in any batch, and the whole solve costs one launch. Needs sys->csr == sys->ilu created
 * with FEMGPU_PLANS_MIXED (rows of at most 16 entries), factor_sweeps 1..16, sweep_mode 1..16,
 * inner_kmax <= 64. threads: 256, 512 or 1024 per workgroup, 0 = chosen from the mean system size.
 * For many small systems (launch latency, not bandwidth, bounds their multi-launch solve). No
 * reference counterpart. */
int femgpu_gmres_mixed_wg(femgpu_ctx *ctx, const femgpu_system *sys, int nsys, const int *row_off,
                          const femgpu_mixed_opts *opt, double tol, int threads, const double *d_F, double *d_X,
                          femgpu_batch_info *info);

/* PCG following Solvers_and_Preconditioners/pcg.c:4-93 (dead code in the reference): stops when
 * gamma < tol^2 gamma0. Returns the iteration count in info->iterations. */
int femgpu_pcg(femgpu_ctx *ctx, const femgpu_system *sys, int itmax, double tol,
               const double *d_B, double *d_X, femgpu_solve_info *info);

/* =========================================================== transport assembly (NDOF = 1) */

typedef struct {
    double k;        /* Condutivity() (kx = ky) */
    double gamma;    /* Reaction() */
    double f0;       /* Font(): constant source */
    double bx[3];    /* Velocity(): Be_x = bx[0] + bx[1]*XB + bx[2]*YB at the element centroid */
    double by[3];    /*             Be_y = by[0] + by[1]*XB + by[2]*YB */
    int shock;       /* 0: no shock capture (Eu = 0); 1: CAU with h_shock = sqrt|2A| */
    int pad;
} femgpu_transport_problem;

typedef struct femgpu_mesh femgpu_mesh;

/* Device-resident mesh for the assembly kernels: node coordinates, Type, invP_id (unknown id in
 * the solve numbering), Dirichlet values ubc (upresc at the node), connectivity ien[nel*3] and
 * element map lm[nel*3] (sink NEQ). Builds the gather maps of the RHS and, when scheme is
 * non-null, of the CSR values (scheme[nel*9] from femgpu_csr_init, nnz slots). */
int femgpu_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x,
                       const double *y, const int *type, const int *invP_id, const double *ubc,
                       const int *ien, const int *lm, const int *scheme, int nnz,
                       femgpu_mesh **out);
/* As femgpu_mesh_create, into an existing transport mesh (its device buffers reused where large
 * enough). No reference counterpart. */
int femgpu_mesh_rebuild(femgpu_mesh *M, int nnodes, int nel, int neq, const double *x, const double *y,
                        const int *type, const int *invP_id, const double *ubc, const int *ien, const int *lm,
                        const int *scheme, int nnz);
int femgpu_mesh_destroy(femgpu_mesh *M);

/* Build_K_F_SUPG (01_SS_Transport_Equation/2D/01_CommonFiles/Build_K_F_SUPG.c:4-187 with
 * F_assembly.c, CAU_ShockCapture.c, h_shock_2sqrtArea.c): element matrices into d_Ke[nel*9]
 * (may be NULL when only CSR values are wanted), RHS into d_F[neq] (F_assembly order), CSR values
 * into d_AA[nnz+1] (csr_assembly order) when d_AA is non-null. d_u: current iterate (neq). */
int femgpu_assemble_transport(femgpu_mesh *M, const femgpu_transport_problem *pb,
                              const double *d_u, double *d_Ke, double *d_F, double *d_AA);

/* Build_M_K_R_SUPG (02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c:6-250 with
 * eval_U_dU.c and its CAU): Ae = Me + alpha*dt*Ke into d_Ae (nullable) / d_AA (nullable),
 * residual into d_R[NEQ+1] (R[NEQ] = 0). d_u, d_du: current u and du (neq). */
int femgpu_assemble_transport_transient(femgpu_mesh *M, const femgpu_transport_problem *pb,
                                        double alpha, double dt, const double *d_u,
                                        const double *d_du, double *d_Ae, double *d_R,
                                        double *d_AA);

/* =========================================================== Navier-Stokes assembly (NDOF = 3) */

/* Device-resident Navier-Stokes mesh: node coordinates, per-dof Type and invP_id (type3, invP3:
 * 3*nnodes, node-major u, v, p; invP_id in the solve numbering), prescribed values presc3
 * (v1presc, v2presc, ppresc at the node), connectivity ien[nel*3], lm[nel*9] (sink NEQ = neq) and,
 * when scheme is non-null, the CSR slot table scheme[nel*81] (femgpu_csr_init with s = 9).
 * Destroy with femgpu_mesh_destroy. */
int femgpu_ns_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x,
                          const double *y, const int *type3, const int *invP3, const double *presc3,
                          const int *ien, const int *lm, const int *scheme, int nnz,
                          femgpu_mesh **out);

/* Build_K_F_SUPG_PSPG (04_SS_NavierStokes_Equations/2D/01_CommonFiles/Build_K_F_SUPG_PSPG.c:8-671,
 * the Picard tangent of :466-566, with eval_U.c:3-38): 9x9 SUPG/PSPG/LSIC element matrices into
 * d_Ke[nel*81] (nullable), residual into d_F[NEQ+1] (F[lm] += Res in element order, F[NEQ] = 0),
 * CSR values into d_AA[nnz+1] (nullable; csr_assembly order). d_u: current iterate (neq);
 * d_fx, d_fy: f1ext/f2ext at the nodes (nullable: zero source); Re: Reynolds number (rho = 1). */
int femgpu_assemble_ns(femgpu_mesh *M, double Re, const double *d_u, const double *d_fx,
                       const double *d_fy, double *d_Ke, double *d_F, double *d_AA);

/* =========================================================== Euler assembly (NDOF = 4) */

typedef struct {
    double gamma;       /* FemFunctions->gamma (constant over the domain) */
    double Mach;        /* reference Mach number (dimensionless Jacobians) */
    double alpha, dt;   /* Alpha_Build, DeltaT_Build */
    double tolerance;   /* AuxBuild->tolerance: CAU acts where |grad U|_A0 >= tolerance */
    int dimensionless;  /* 1: dimensionless_Ax_Ay_calculations, 0: dimensional_Ax_Ay_calculations */
    int shock;          /* 1: Delta_CAU, 0: no shock capture */
} femgpu_euler_problem;

/* Device-resident Euler mesh: as femgpu_ns_mesh_create with 4 dofs per node (rho, rho u, rho v,
 * rho E; presc4 = rhopresc, v1presc, v2presc, epresc at the node), lm[nel*12] and
 * scheme[nel*144] (femgpu_csr_init with s = 12). */
int femgpu_euler_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x,
                             const double *y, const int *type4, const int *invP4,
                             const double *presc4, const int *ien, const int *lm,
                             const int *scheme, int nnz, femgpu_mesh **out);

/* Build_M_K_F_SUPG (03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c:6-1057 with
 * eval_U_dU.c, Ax_Ay_calculations.c and Delta_CAU.c): Ae = Me + alpha*dt*Ke (12x12) into
 * d_Ae[nel*144] (nullable), R[lm] += -(Me dU + Ke U) into d_R[NEQ+1] (R[NEQ] = 0), CSR values into
 * d_AA[nnz+1] (nullable). d_u, d_du: current state and increment (neq). */
int femgpu_assemble_euler(femgpu_mesh *M, const femgpu_euler_problem *pb, const double *d_u,
                          const double *d_du, double *d_Ae, double *d_R, double *d_AA);

/* =========================================================== nodal block-diagonal preconditioner */

/* BlockDiag_precond_EBE_setup / BlockDiagDOF3_precond_EBE_setup (00_CommonFiles/
 * Solvers_and_Preconditioners/EBE/BlockDiag_precond_EBE_setup.c:6-88, EBE/BlockDiagDOF3_precond_EBE_setup.c)
 * for ndof = 4 / 3 (1: point Jacobi): the node -> dof equations (type/invP per dof, nnodes*ndof;
 * prescribed dofs map to the sink NEQ) and the node -> element incidences of ien[nel*3]. */
int femgpu_blockdiag_create(femgpu_ctx *ctx, int nnodes, int nel, int ndof, int NEQ, const int *ien,
                            const int *type, const int *invP, femgpu_blockdiag **out);
int femgpu_blockdiag_destroy(femgpu_blockdiag *B);
/* The ndof x ndof diagonal block of every node summed over its elements in element order
 * (BlockDiag_precond_EBE_setup.c:66-77) from d_Ae (nel*s*s, s = 3*ndof, element order, the layout
 * of the assembly kernels), inverted (:79-80). The reference then applies it to F (:84): call
 * femgpu_blockdiag_apply(B, F, F). */
int femgpu_blockdiag_setup(femgpu_blockdiag *B, const double *d_Ae);
/* BlockDiag_precond / BlockDiagDOF3_precond (BlockDiag_precond.c:3-29): z = D^-1 p per node,
 * prescribed dofs read as 0 and left unwritten. d_p == d_z allowed. */
int femgpu_blockdiag_apply(femgpu_blockdiag *B, const double *d_p, double *d_z);
/* Block_scaling_EBE (Block_scaling.c:10-118): blocks from the unscaled d_Ae, F := D^-1 F (d_F
 * nullable), then every element row block a := D_{v_a}^-1 (row block with the rows and columns of
 * prescribed dofs zeroed and a unit diagonal), in place. */
int femgpu_block_scaling(femgpu_blockdiag *B, double *d_Ae, double *d_F);

/* =========================================================== meshes and mesh files (host) */

/* The structured generator of Mesh_Generator/Structured_Mesh/meshGenerator.c:95-190: (M+1)(N+1)
 * nodes row by row on [xinf, xsup] x [yinf, ysup], 2MN triangles, two per square, squares row by
 * row, cut along the lower-left/upper-right diagonal when right = 1. x, y: (M+1)(N+1); ien: 6MN. */
int femgpu_structured_mesh(int M, int N, double xinf, double yinf, double xsup, double ysup, int right,
                           double *x, double *y, int *ien);

/* setmark_according_boundary (meshGenerator.c:221-363) for the named problem (OBLIQUO, REFLETIDO,
 * EXPLOSION, SOD, BAROCLINIC, PUDIM, CONE, CARTOLA, TESTE, CAVITY): the boundary mark of every node
 * and (type non-null: nnodes*NDOF) the Type of every dof from the mark table (:16-19). Returns NDOF
 * (1, 3 or 4) or a negative status for an unknown problem. */
int femgpu_structured_marks(const char *problem, int nnodes, const double *x, const double *y, int *mark,
                            int *type);

typedef struct {
    int nnodes, nel, ndof, format;
} femgpu_mesh_file_info;

/* Reads a mesh file. format 0: the generator's `<PROBLEM>_<nn>_<ne>.dat` (meshGenerator.c:193-211);
 * format 1: one rank's `RANK_<r>_<PROBLEM>_<nn>_<ne>.dat` as the drivers read it
 * (01_SS_Transport_Equation/2D/01_CommonFiles/Preprocess.c:127-168: per dof "0" or
 * "Type Send invP_id"). Call with null arrays for the sizes, then with x, y (nnodes), type and invP
 * (nnodes*ndof; invP -1 where the file has none), ien (3*nel), etype (nel; -1 where absent). */
int femgpu_mesh_file_read(const char *path, int format, int ndof, femgpu_mesh_file_info *info, double *x,
                          double *y, int *type, int *invP, int *ien, int *etype);

/* Writes the generator's format (meshGenerator.c:193-211), byte for byte as the generator prints. */
int femgpu_mesh_file_write(const char *path, int nnodes, int nel, int ndof, const double *x, const double *y,
                           const int *type, const int *ien);

/* Delaunay triangulation of n points (host arrays), for unstructured meshes (the reference makes
 * them with gmsh, MG/Unstructured_Mesh/gmsh2Mesh.c, which is not in its tree): Bowyer-Watson with
 * ghost triangles over the hull, Hilbert-ordered insertion. tri: 3 * max_tri ints, counter-
 * clockwise triangles (at most 2n - 5); points on hull edges are vertices; repeated points are
 * skipped and counted in *duplicates (may be null). Fails if all points are collinear. */
int femgpu_delaunay(int n, const double *x, const double *y, int max_tri, int *ntri, int *tri, int *duplicates);

/* =========================================================== serving loop */

typedef struct femgpu_server femgpu_server;

typedef struct {
    size_t hbm_budget;     /* device bytes the work in flight may hold (estimated per request) */
    int executors;         /* concurrent HIP streams, each with its own library context (1..8) */
    int prep_threads;      /* host threads preparing requests (numbering, lm, SYMRCM CSR) */
    int small_rows;        /* requests with at most this many unknowns are packed into batches */
    int batch_rows;        /* unknowns of one batch at most */
    int max_batch;         /* requests per batch; 0: tuned online among 4, 8, 16, 32, 64 */
    int kmax;              /* inner GMRES basis size */
    int small_solver;      /* batches of small requests: 0 the block-diagonal shared-Krylov solver
                              (femgpu_gmres_mixed_batched), 1 the persistent launch with a workgroup
                              per system (femgpu_gmres_mixed_wg) when every system of the batch has
                              at most 3072 rows, else 0, 2 always the persistent launch; the
                              persistent launch needs kmax <= 64 */
    int wg_threads;        /* threads per workgroup of the persistent launch (256, 512, 1024); 0:
                              tuned online among the three once the batch size is settled */
    double tol;            /* FP64-verified relative residual of every request */
    femgpu_mixed_opts mixed;   /* FP32 preconditioner (inner_kmax is taken from kmax) */
} femgpu_server_opts;

typedef struct {
    int n, nnz;            /* unknowns and CSR entries of the request's system */
    int batch_id, batch_size;
    int iterations;        /* FP32 Arnoldi steps the request took part in */
    int refinements;       /* FP64 refinement steps */
    int status;            /* 0 converged, 1 refinement cap, < 0 an error status */
    int wg_threads;        /* threads per workgroup when solved by the persistent launch, else 0 */
    double true_rel_residual;      /* ||F - A x|| / ||F|| in FP64 */
    double t_submit, t_start, t_end;   /* seconds on the server clock */
    double prep_s;         /* host preparation of the request (numbering, lm, CSR + SYMRCM) */
    double setup_s;        /* its unit's device set-up: CSR plans, meshes, assembly */
    double solve_s;        /* its unit's batched solve and result download */
} femgpu_request_info;

/* A request loop for steady transport solves (Build_K_F_SUPG + mixed-precision GMRES with the
 * parallel ILU(0) / Jacobi preconditioner) on meshes of any size, with continuous batching: host
 * threads prepare requests while executors solve; a large request runs alone, consecutive small
 * ones are packed into one block-diagonal batched solve (femgpu_gmres_mixed_batched); admission
 * keeps the estimated device memory of the work in flight under hbm_budget. No reference
 * counterpart. */
int femgpu_server_create(int device, const femgpu_server_opts *opt, femgpu_server **out);
/* Queue one request (host arrays copied: nodes x, y, Type 0/1, Dirichlet values ubc; ien[nel*3];
 * the problem). Thread-safe. */
int femgpu_server_submit(femgpu_server *S, int nnodes, int nel, const double *x, const double *y,
                         const int *type, const double *ubc, const int *ien,
                         const femgpu_transport_problem *pb, long long *ticket);
/* Wait up to timeout_s for a request: returns FEMGPU_OK with the nodal solution u[nnodes] and
 * its info (the ticket is then released), or 1 on timeout. */
int femgpu_server_wait(femgpu_server *S, long long ticket, double timeout_s, double *u,
                       femgpu_request_info *info);
/* Seconds since the server started (the clock of the info timestamps). */
double femgpu_server_clock(femgpu_server *S);
/* The batch size in use (0 while the online tuning is still sampling). */
int femgpu_server_batch_tuning(femgpu_server *S, int *chosen);
/* The persistent launch's threads per workgroup in use (0 while still sampling). */
int femgpu_server_wg_tuning(femgpu_server *S, int *chosen);
/* Stops accepting requests, finishes the queued ones and joins the threads. */
int femgpu_server_destroy(femgpu_server *S);

/* =========================================================== nonlinear and time drivers */

typedef struct {
    int solver;            /* 0: ILUp_precond_setup + pgmres (FP64, the reference's pair);
                              1: femgpu_gmres_mixed on the raw right-hand side */
    int kmax, lmax;        /* KrylovBasisVectorsQuantity, LinearMaxIter */
    double tol;            /* SolverTolerance */
    femgpu_mixed_opts mixed;   /* solver = 1 */
    double nonlinear_tol;  /* NonLinearTolerance */
    int nonlinear_maxiter; /* NonLinearMaxIter */
    int stop_by_norm;      /* time step: 1 StopByNorm, 0 StopByIterations (Criteria.c:25-41) */
} femgpu_driver_opts;

typedef struct {
    int iterations;        /* fixed-point / Picard iterations, or multicorrections of the step */
    int linear_iterations; /* Arnoldi steps summed over the linear solves */
    int status;            /* 0 stopped by the criterion, 1 by the iteration cap, 2 a linear solve
                              did not converge */
    int pad;
    double normdiff, normu;    /* steady: last ||u - uold|| (NS: ||s||) and ||u|| */
    double norm_a, norm_Da;    /* time step: last ||a|| and ||Da|| */
    int iters[64];         /* Arnoldi steps of each linear solve (the first 64) */
} femgpu_driver_info;

/* Steady transport fixed point (01_SS_Transport_Equation/2D/01_CommonFiles/Process.c:29-86):
 * repeat Build_K_F_SUPG at u (CSR values assembled into A), solve, until
 * ||u - uold|| <= NonLinearTolerance ||u|| or nonlinear_maxiter iterations. Scaling "NOT" only.
 * d_u (neq): initial iterate in, final solution out. M must carry A's slot table. */
int femgpu_st_process(femgpu_mesh *M, femgpu_csr *A, const femgpu_transport_problem *pb,
                      const femgpu_driver_opts *opt, double *d_u, femgpu_driver_info *info);

/* Navier-Stokes Picard loop (04_SS_NavierStokes_Equations/2D/01_CommonFiles/Process.c:27-52):
 * repeat Build_K_F_SUPG_PSPG at u, solve for s, u += s, until ||s|| / ||u|| <= NonLinearTolerance
 * or nonlinear_maxiter iterations. d_fx, d_fy as femgpu_assemble_ns. */
int femgpu_ns_process(femgpu_mesh *M, femgpu_csr *A, double Re, const double *d_fx, const double *d_fy,
                      const femgpu_driver_opts *opt, double *d_u, femgpu_driver_info *info);

/* One step of the predictor-multicorrector (00_CommonFiles/Time_Integration/Predictor_Old.c:53-123)
 * for the transient transport: u_old = u, u += (1-alpha) dt a, a = 0, then multicorrections
 * (Build_M_K_R_SUPG at (u, a), solve for Da, a += Da, u += alpha dt Da) until the stop criterion.
 * d_u, d_a (neq): the state, updated in place; the caller advances t. */
int femgpu_tt_predictor_old_step(femgpu_mesh *M, femgpu_csr *A, const femgpu_transport_problem *pb,
                                 double dt, double alpha, const femgpu_driver_opts *opt, double *d_u,
                                 double *d_a, femgpu_driver_info *info);

/* One step of the predictor-multicorrector for the Euler equations (03_Euler_Equations/2D/
 * 01_CommonFiles/Process.c:4-21 -> 00_CommonFiles/Time_Integration/Predictor_Old.c:53-123) with the
 * EBE scheme and the BlockDiag preconditioner: u_old = u, u += (1-alpha) dt a, a = 0, then
 * multicorrections -- Build_M_K_F_SUPG at (u, a) (element matrices into E, residual R), the nodal
 * block-diagonal blocks (BlockDiag_precond_EBE_setup), and either R := D^-1 R and pgmres with
 * ebemvNDOF4 and BlockDiag_precond (solver 0) or the mixed solver on the raw R with FP32 element
 * matrices and blocks (solver 1) for Da; a += Da, u += alpha dt Da -- until the stop criterion.
 * alpha and dt are pb's. E: femgpu_ebe_create(s = 12) on M's lm; B: femgpu_blockdiag_create(ndof 4)
 * on M's nodes. Process.c sets a scaling function (setScaling.c) that Predictor_Old never calls,
 * so none is applied. d_u, d_a (neq): the state, updated in place. */
int femgpu_eu_predictor_old_step(femgpu_mesh *M, femgpu_ebe *E, femgpu_blockdiag *B,
                                 const femgpu_euler_problem *pb, const femgpu_driver_opts *opt, double *d_u,
                                 double *d_a, femgpu_driver_info *info);

#ifdef __cplusplus
}
#endif
#endif
