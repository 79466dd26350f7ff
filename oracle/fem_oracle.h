/*
 * fem_oracle.h -- CPU FP64 restatement of the PARALLEL_FEM_CODES linear-solve hot path.
 *
 * THIS IS TEST INFRASTRUCTURE. It is the checker for the HIP library, not part of the product:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 * The product path (libfemgpu.so) never links, calls or falls back to it.
 *
 * Every function restates the reference algorithm from its source (file:line cited beside each
 * declaration, paths relative to the reference root) with plain IEEE FP64 arithmetic
 * (compiled with -ffp-contract=off, no fast-math) and the reference's own loop and reduction
 * orders, so results can be compared bit-for-bit with a harness build of the reference sources.
 *
 * Conventions: one rank (neq_bef = 0, NEQ = neq), 0-based indices, flat row-major arrays.
 * lm[e*s + a] in [0, NEQ]; NEQ is the Dirichlet sink. The CSR matrix of neq rows has nnz entries
 * and value arrays of length nnz+1 (slot nnz is the dummy written by Dirichlet couplings).
 */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ mesh + numbering */

/* Structured M x N squares, each cut on one diagonal; node k = j*(M+1)+i at
 * (xinf + i*hx, yinf + j*hy). right_cut=1: "right" (MG/Structured_Mesh/meshGenerator.c:113-138),
 * 0: "left" (:140-166). x,y: (M+1)(N+1); ien: 2MN*3. */
void orc_mesh_structured(int M, int N, double xinf, double yinf, double xsup, double ysup,
                         int right_cut, double *x, double *y, int *ien);

/* Global unknown numbering as Parallel_Preprocess writes it with reordering NOT and one part
 * (PP/Preprocess.c:37-53, :299-305): unknown id = running count of Type-1 dofs in node order,
 * -1 for Dirichlet dofs. type/invP_id are nnodes*ndof (node-major). Returns neq. */
int orc_number_unknowns(int nnodes, int ndof, const int *type, int *invP_id);

/* Fill_LM (01_SS_Transport_Equation/2D/01_CommonFiles/Fill_LM.c:3-27 for ndof=1,
 * 04_SS_NavierStokes_Equations/2D/01_CommonFiles/Fill_LM.c:3-31 for ndof>1):
 * lm[e*(3*ndof) + a*ndof + d] = invP_id of dof d of vertex a, or NEQ when Dirichlet. */
void orc_fill_lm(int nel, int ndof, int NEQ, const int *ien, const int *type, const int *invP_id,
                 int *lm);

/* ------------------------------------------------------------------ reordering (integer, bit-exact) */

/* REORDERING_SYMRCM (00_CommonFiles/Reordering/symrcm.c:4-237, find_starting_node :352-483,
 * heap :505-567) on a structurally symmetric CSR pattern. perm[new] = old. */
void orc_symrcm(int n, const int *IA, const int *JA, int *perm);

/* Row then column permutation with per-row re-sort, carrying slot ids in pT
 * (00_CommonFiles/Reordering/reordering.c:101-167). perm[new] = old. In place. */
void orc_csr_permute(int n, int nnz, int *IA, int *JA, const int *perm, int *pT);

/* ------------------------------------------------------------------ CSR structure (csr_Initialization) */

/* Number of nonzeros of the local block built from lm (csr_Initialization.c:36-99). */
int orc_csr_nnz(int nel, int s, int neq, const int *lm);

/* csr_Initialization for one rank (01_SS_Transport_Equation/2D/01_CommonFiles/csr_Initialization.c:5-571):
 * sorted-row CSR from lm, per-element slot table scheme[e*s*s + a*s + b] (nnz = dummy),
 * reordering (reorder: 0 = NOT, 1 = SYMRCM; Spectral is third-party and not provided),
 * remap of slots and of invP_id (Type-1 dofs only). When permute_lm != 0 the element map
 * lm is renumbered too (02_Transport_Equation/2D/01_CommonFiles/csr_Initialization.c:452-454).
 * perm[new] = old. Returns nnz, or -1 on an unsupported reorder code. */
int orc_csr_init(int nel, int s, int neq, int *lm, int reorder, int permute_lm,
                 int ndofs_total, const int *type, int *invP_id,
                 int *IA, int *JA, int *scheme, int *perm);

/* Permutation_of_LM (01_SS_Transport_Equation/2D/01_CommonFiles/Permutation_of_LM.c:37-105):
 * node adjacency of unknowns from the elements, reordered; Perm/invPerm have NEQ+1 entries and
 * Perm[NEQ] = invPerm[NEQ] = NEQ. Returns the adjacency nnz (Parameters->nnzero there). */
int orc_ebe_permutation(int nel, int NEQ, const int *ien, const int *type, const int *invP_id,
                        int reorder, int *Perm, int *invPerm);

/* ------------------------------------------------------------------ transport problem (NDOF = 1) */

typedef struct {
    double k;        /* Condutivity() (kx = ky) */
    double gamma;    /* Reaction() */
    double f0;       /* Font(): constant source */
    double bx[3];    /* Velocity(): Be_x = bx[0] + bx[1]*XB + bx[2]*YB at the centroid */
    double by[3];    /*             Be_y = by[0] + by[1]*XB + by[2]*YB */
    int shock;       /* 0: Eu = 0 (linear); 1: CAU with h_shock = sqrt|2A| */
    int pad;
} orc_transport_problem;

/* Build_K_F_SUPG (01_SS_Transport_Equation/2D/01_CommonFiles/Build_K_F_SUPG.c:4-187) with
 * F_assembly (F_assembly.c:3-28), CAU_ShockCapture (CAU_ShockCapture.c:3-31) and
 * h_shock_2sqrtArea (h_shock_2sqrtArea.c:3-7). ubc[node] = upresc(x,y) for Dirichlet nodes;
 * u[neq] is the current iterate. Writes Ke[nel*9] (ebe_assembly.c layout) and F[neq]. */
void orc_build_transport_supg(int nnodes, int nel, int neq, const double *x, const double *y,
                              const int *type, const int *invP_id, const int *ien,
                              const double *ubc, const double *u, const orc_transport_problem *pb,
                              double *Ke, double *F);

/* Build_M_K_R_SUPG (02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c:6-250) with
 * eval_U_dU (eval_U_dU.c:3-38) and the transient CAU (CAU_ShockCapture.c:3-31 there).
 * Ae = Me + alpha*dt*Ke per element into Ke[nel*9]; R[NEQ+1] residual indexed through lm. */
void orc_build_transport_transient(int nnodes, int nel, int NEQ, const double *x, const double *y,
                                   const int *type, const int *invP_id, const int *ien,
                                   const int *lm, const double *ubc, const double *u,
                                   const double *du, const orc_transport_problem *pb,
                                   double alpha, double dt, double *Ae, double *R);

/* eval_U (04_SS_NavierStokes_Equations/2D/01_CommonFiles/eval_U.c:3-38): nodal state
 * U3[3I+d] (d = u, v, p) = u[invP3[3I+d]] when type3[3I+d] > 0, else the prescribed presc3[3I+d]. */
void orc_ns_nodal_state(int nnodes, const int *type3, const int *invP3, const double *u,
                        const double *presc3, double *U3);

/* Build_K_F_SUPG_PSPG (04_SS_NavierStokes_Equations/2D/01_CommonFiles/Build_K_F_SUPG_PSPG.c:8-671),
 * the Picard tangent (the branch taken while its call counter is below 50000, :466-566): per element
 * the 9x9 SUPG/PSPG/LSIC matrix (dofs u, v, p per vertex) into Ke[nel*81] row-major
 * (ebe_assembly.c:5-17) and the residual Res added to F[lm] in element order, F[NEQ] = 0
 * (:655-658). U3: nodal state (orc_ns_nodal_state); fx, fy: f1ext/f2ext at the nodes; rho = 1,
 * viscosity 1/Re (:120-125). */
void orc_build_ns_supg_pspg(int nnodes, int nel, int NEQ, const double *x, const double *y,
                            const int *ien, const int *lm, const double *U3, const double *fx,
                            const double *fy, double Re, double *Ke, double *F);

/* ------------------------------------------------------------------ Euler (NDOF = 4) */

typedef struct {
    double gamma;          /* FemFunctions->gamma (constant over the domain) */
    double Mach;           /* reference Mach number (dimensionless Jacobians only) */
    double alpha, dt;      /* Alpha_Build, DeltaT_Build */
    double tolerance;      /* AuxBuild->tolerance of the shock-capture switch */
    int dimensionless;     /* 1: dimensionless_Ax_Ay_calculations, 0: dimensional (Ax_Ay_calculations.c) */
    int shock;             /* 1: Delta_CAU (Delta_CAU.c:4-83), 0: delta = 0 */
} orc_euler_params;

/* eval_U_dU (03_Euler_Equations/2D/01_CommonFiles/eval_U_dU.c:3-55): nodal (rho, rho u, rho v,
 * rho E) and their increments; prescribed dofs take presc4 and a zero increment. */
void orc_euler_nodal_state(int nnodes, const int *type4, const int *invP4, const double *u,
                           const double *du, const double *presc4, double *U4, double *dU4);

/* Build_M_K_F_SUPG (03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c:6-1057): per element
 * Ae = Me + alpha*dt*Ke (12x12, dofs rho, rho u, rho v, rho E per vertex) into Ae[nel*144] and
 * R[lm] += -(Me dUe + Ke Ue) in element order, R[NEQ] = 0. */
void orc_build_euler_supg(int nnodes, int nel, int NEQ, const double *x, const double *y,
                          const int *ien, const int *lm, const double *U4, const double *dU4,
                          const orc_euler_params *p, double *Ae, double *R);

/* csr_assembly (csr_assembly.c:3-54, one rank): AA[nnz+1] zeroed then AA[scheme] += Ke per
 * element in element order; AA[nnz] = 0. */
void orc_csr_assemble(int nel, int s, int nnz, const int *scheme, const double *Ke, double *AA);

/* ------------------------------------------------------------------ operators */

/* ebemv (00_CommonFiles/MatrixVector_Operations/EBE/ebemv.c:4-64), NDOF = 1 with the local
 * permutation; P and Q have NEQ+1 entries, P[NEQ] is read as 0 through the permuted copy. */
void orc_ebemv(int nel, int NEQ, const int *lm, const int *Perm, const int *invPerm,
               const double *A, const double *P, double *Q);

/* Block EBE matvec for s x s element blocks (ebemvNDOF3.c:3-92 s=9, ebemvNDOF4.c:4-121 s=12):
 * Q[0..NEQ) zeroed, Q[lm[a]] += sum_b A[e][a*s+b] P[lm[b]] with P[NEQ] treated as 0. */
void orc_ebemv_block(int nel, int s, int NEQ, const int *lm, const double *A, const double *P,
                     double *Q);

/* csrmv (00_CommonFiles/MatrixVector_Operations/CSR/csrmv.c:4-65), local block only. */
void orc_csrmv(int n, const int *IA, const int *JA, const double *AA, const double *P, double *Q);

/* ------------------------------------------------------------------ BLAS (00_CommonFiles/BLAS_Operations) */

double orc_ddot(int n, const double *x, const double *y);      /* ddot.c:4-21, 5-way shape */
void   orc_daxpy(int n, double a, const double *x, double *y); /* daxpy.c:3-20 */
void   orc_dscal(int n, double a, double *x);                  /* dscal.c:3-20 */

/* ------------------------------------------------------------------ ILU(0) */

/* ILUp_precond_setup numeric part for p = 0 (CSR/ILUp_precond_setup.c:67-168 with the level-0
 * pattern of ilup.c:164-302, which for p = 0 is A's own pattern, L = strict lower part in
 * ascending column order, U = strict upper part in CSR order). LU[nnz] holds the factors in
 * A's layout (L multipliers below, U above the diagonal); Dinv[n] = 1/pivot. Returns 0, or
 * 1 + row of the first zero pivot / missing diagonal (the reference exits there, :155-164). */
int orc_ilu0_factor(int n, const int *IA, const int *JA, const double *AA, double *LU,
                    double *Dinv);

/* ILUp_precond (CSR/ILUp_precond.c:4-47): z = U^-1 L^-1 p, p == z allowed. */
void orc_ilu0_apply(int n, const int *IA, const int *JA, const double *LU, const double *Dinv,
                    const double *p, double *z);

/* ------------------------------------------------------------------ nodal block diagonal */

/* BlockDiag_precond_EBE_setup / BlockDiagDOF3_precond_EBE_setup numeric part
 * (00_CommonFiles/Solvers_and_Preconditioners/EBE/BlockDiag_precond_EBE_setup.c:66-80 and the DOF3
 * file): per node the ndof x ndof diagonal block of its elements' matrices A[nel*s*s] (s = 3 ndof,
 * element order), inverted by the reference's cofactor formulas into inv[nnodes*ndof*ndof]. */
void orc_blockdiag_setup(int nnodes, int nel, int ndof, const int *ien, const double *A, double *inv);

/* BlockDiag_precond / BlockDiagDOF3_precond (BlockDiag_precond.c:3-29): z[id] = inv p[id] per
 * node, id[nnodes*ndof] = unknown or NEQ (prescribed, read as 0, not written); p == z allowed. */
void orc_blockdiag_apply(int nnodes, int ndof, int NEQ, const int *id, const double *inv, const double *p,
                         double *z);

/* ------------------------------------------------------------------ solvers */

/* Operator + preconditioner selection for the oracle's pgmres. */
typedef struct {
    int kind;              /* 0: CSR (IA,JA,AA), 1: EBE NDOF=1 (lm,Perm,invPerm,A), 2: block EBE */
    int n;                 /* neq */
    int nel, s;
    const int *IA, *JA;    /* CSR operator */
    const double *AA;
    const int *lm, *Perm, *invPerm;  /* EBE operator */
    const double *A;
    int precond;           /* 0: NOT, 1: ILU(0) from (pIA,pJA,LU,Dinv), 2: nodal block diagonal */
    const int *pIA, *pJA;
    const double *LU, *Dinv;
    int bd_nnodes, bd_ndof;  /* precond 2: orc_blockdiag_apply(bd_nnodes, bd_ndof, n, bd_id, bd_inv) */
    const int *bd_id;
    const double *bd_inv;
} orc_system;

/* pgmres (00_CommonFiles/Solvers_and_Preconditioners/pgmres.c:5-200): restarted, left
 * preconditioned, MGS, Givens; kmax = KrylovBasisVectorsQuantity (kmax-1 Arnoldi steps per
 * cycle), lmax = LinearMaxIter cycles, X zeroed on entry. B is the already preconditioned RHS.
 * Returns the number of Arnoldi steps (added to Parameters->iterations there). B and X have
 * n+1 entries. *rho_out receives the final Givens residual estimate. */
int orc_pgmres(const orc_system *sys, int kmax, int lmax, double tol, const double *B, double *X,
               double *rho_out, int *cycles_out);

/* PCG as written in Solvers_and_Preconditioners/pcg.c:4-93 (dead code in the reference):
 * stop when gamma < tol^2 gamma0; returns the iteration count it stores. */
int orc_pcg(const orc_system *sys, int itmax, double tol, const double *B, double *X);

#ifdef __cplusplus
}
#endif
#endif
