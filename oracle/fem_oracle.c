/*
 * fem_oracle.c -- CPU FP64 restatement of the PARALLEL_FEM_CODES linear-solve hot path.
 *
 * TEST INFRASTRUCTURE (the checker): loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. Never linked into, called by, or used as a fallback of the
 * HIP product library.
 *
 * Parity status per function (how each is pinned, see the tests named in brackets):
 *   orc_mesh_structured, orc_number_unknowns, orc_fill_lm, orc_symrcm, orc_csr_permute,
 *   orc_csr_nnz, orc_csr_init, orc_ebe_permutation       -- bit-exact vs reference harness
 *                                                            golden fixtures [test_oracle_golden]
 *   orc_build_transport_supg, orc_csr_assemble           -- golden fixtures + TESTE/patch
 *                                                            identities [test_oracle_golden,
 *                                                            test_oracle_math]
 *   orc_build_transport_transient                        -- the steady fixtures' stiffness plus the
 *                                                            textbook SUPG mass with du != 0, and the
 *                                                            transient CAU residual [test_oracle_math]
 *   orc_blockdiag_setup, orc_blockdiag_apply             -- dense inverse of the summed nodal
 *                                                            blocks [test_oracle_math]
 *   orc_ebemv, orc_ebemv_block, orc_csrmv                 -- golden fixtures + EBE == CSR +
 *                                                            dense brute force [test_oracle_math]
 *   orc_ilu0_factor, orc_ilu0_apply                      -- golden fixtures + exact LU on
 *                                                            elimination-closed patterns
 *   orc_pgmres                                           -- golden iteration counts/solutions +
 *                                                            d-distinct-eigenvalue termination
 *   orc_pcg                                              -- CG invariants only (the reference
 *                                                            never calls pcg.c)
 * Compile with -O2 -ffp-contract=off (no fast-math): every loop and reduction keeps the
 * reference's order so results match a strict-IEEE build of the reference sources.
 */
#include "fem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== mesh + numbering */

void orc_mesh_structured(int M, int N, double xinf, double yinf, double xsup, double ysup,
                         int right_cut, double *x, double *y, int *ien)
{
    /* Coordinates: meshGenerator.c:94-96 (hx, hy) and :170-177 (node k = j*nosx + i). */
    const int nosx = M + 1, nosy = N + 1;
    const double hx = (xsup - xinf) / M;
    const double hy = (ysup - yinf) / N;
    int i, j, k = 0;
    for (j = 0; j < nosy; j++)
        for (i = 0; i < nosx; i++, k++) {
            x[k] = xinf + i * hx;
            y[k] = yinf + j * hy;
        }
    /* Connectivity: squares row by row, two triangles each. The "right" cut joins the lower-left
     * corner to the upper-right one (meshGenerator.c:113-138), the "left" cut the lower-right to
     * the upper-left (:140-166); vertex order inside each triangle follows those loops. */
    int e = 0;
    for (j = 0; j < N; j++)
        for (i = 0; i < M; i++) {
            const int a = j * nosx + i;
            if (right_cut) {
                ien[3 * e + 0] = a;     ien[3 * e + 1] = a + 1;        ien[3 * e + 2] = a + 1 + nosx; e++;
                ien[3 * e + 0] = a;     ien[3 * e + 1] = a + 1 + nosx; ien[3 * e + 2] = a + nosx;     e++;
            } else {
                ien[3 * e + 0] = a;     ien[3 * e + 1] = a + 1;        ien[3 * e + 2] = a + nosx;     e++;
                ien[3 * e + 0] = a + 1; ien[3 * e + 1] = a + 1 + nosx; ien[3 * e + 2] = a + nosx;     e++;
            }
        }
}

int orc_number_unknowns(int nnodes, int ndof, const int *type, int *invP_id)
{
    int I, d, neq = 0;
    for (I = 0; I < nnodes; I++)
        for (d = 0; d < ndof; d++)
            invP_id[I * ndof + d] = (type[I * ndof + d] == 1) ? neq++ : -1;
    return neq;
}

void orc_fill_lm(int nel, int ndof, int NEQ, const int *ien, const int *type, const int *invP_id,
                 int *lm)
{
    const int s = 3 * ndof;
    int e, a, d;
    for (e = 0; e < nel; e++)
        for (a = 0; a < 3; a++) {
            const int J = ien[3 * e + a];
            for (d = 0; d < ndof; d++)
                lm[e * s + a * ndof + d] = (type[J * ndof + d] != 0) ? invP_id[J * ndof + d] : NEQ;
        }
}

/* ====================================================================== SYMRCM */

/* A BFS record: node id, degree, distance to the BFS root. */
typedef struct { int id, deg, dist; } rcm_rec;

/* Min-heap on deg with strict comparisons; the exact sift rules decide ties and therefore the
 * permutation, so they follow symrcm.c:505-567 (insert = sift-up while child < parent;
 * remove-min = move last to root, then sift-down preferring left, then right, on strict <). */
static void heap_push(rcm_rec *H, int *h, rcm_rec o)
{
    int i = (*h)++;
    H[i] = o;
    while (i > 0) {
        const int p = (i - 1) >> 1;
        if (H[i].deg < H[p].deg) {
            rcm_rec t = H[i]; H[i] = H[p]; H[p] = t;
            i = p;
        } else
            break;
    }
}

static rcm_rec heap_pop(rcm_rec *H, int *h)
{
    rcm_rec r = H[0];
    const int size = --(*h);
    H[0] = H[size];
    int j = 0;
    for (;;) {
        const int l = 2 * j + 1, rr = 2 * j + 2;
        int m = (l < size && H[l].deg < H[j].deg) ? l : j;
        if (rr < size && H[rr].deg < H[m].deg) m = rr;
        if (m == j) break;
        rcm_rec t = H[j]; H[j] = H[m]; H[m] = t;
        j = m;
    }
    return r;
}

/* Neighbour walk used by both BFS passes of symrcm.c (:116-176 and :386-459): the column of i
 * (CSR row i, as the reference reads the CSR arrays as CCS) merged with the row of i from the
 * transpose, ascending, visiting an index present in both once. visit_cb is applied in that order. */
typedef struct {
    const int *c1, *r1, *c2, *r2;
} rcm_graph;

static int rcm_neigh(const rcm_graph *g, int i, int *out)
{
    int j1 = g->c1[i], j2 = g->c2[i], n = 0;
    const int e1 = g->c1[i + 1], e2 = g->c2[i + 1];
    while (j1 < e1 || j2 < e2) {
        if (j1 == e1) out[n++] = g->r2[j2++];
        else if (j2 == e2) out[n++] = g->r1[j1++];
        else {
            const int a = g->r1[j1], b = g->r2[j2];
            if (a <= b) { out[n++] = a; j1++; if (a == b) j2++; }
            else { out[n++] = b; j2++; }
        }
    }
    return n;
}

/* find_starting_node (symrcm.c:352-483): repeated BFS to a pseudo-peripheral node. */
static int rcm_start(int N, const rcm_graph *g, const int *D, int start, rcm_rec *Q, int *visit,
                     int *nb)
{
    int qh = 0, qt = 0, i;
    rcm_rec x = {start, D[start], 0};
    memset(visit, 0, sizeof(int) * (size_t)N);
    Q[qt] = x; qt = (qt + 1) % (N + 1);
    visit[start] = 1;
    int max_dist = 0;
    for (;;) {
        while (qh != qt) {
            rcm_rec v = Q[qh]; qh = (qh + 1) % (N + 1);
            if (v.dist > x.dist || (v.id != x.id && v.deg > x.deg)) x = v;
            const int nn = rcm_neigh(g, v.id, nb);
            for (i = 0; i < nn; i++) {
                const int r = nb[i];
                if (!visit[r]) {
                    rcm_rec w = {r, D[r], v.dist + 1};
                    Q[qt] = w; qt = (qt + 1) % (N + 1);
                    visit[r] = 1;
                }
            }
        }
        if (max_dist < x.dist) {
            max_dist = x.dist;
            memset(visit, 0, sizeof(int) * (size_t)N);
            visit[x.id] = 1;
            x.dist = 0;
            qt = qh = 0;
            Q[qt] = x; qt = (qt + 1) % (N + 1);
        } else
            break;
    }
    return x.id;
}

void orc_symrcm(int N, const int *IA, const int *JA, int *P)
{
    int i, j, k, l;
    const int nnz = IA[N];
    /* transpose (symrcm.c:272-303) */
    int *c2 = (int *)calloc((size_t)N + 1, sizeof(int));
    int *r2 = (int *)calloc((size_t)nnz + 1, sizeof(int));
    int *w = (int *)calloc((size_t)N + 1, sizeof(int));
    for (i = 0; i < nnz; i++) w[JA[i]]++;
    int acc = 0;
    for (i = 0; i < N; i++) { c2[i] = acc; acc += w[i]; w[i] = c2[i]; }
    c2[N] = acc;
    for (j = 0; j < N; j++)
        for (k = IA[j]; k < IA[j + 1]; k++) r2[w[JA[k]]++] = j;
    free(w);
    /* calc_degrees (symrcm.c:305-350) */
    int *D = (int *)calloc((size_t)N, sizeof(int));
    int max_deg = 0;
    for (j = 0; j < N; j++)
        for (i = IA[j]; i < IA[j + 1]; i++) {
            k = JA[i];
            D[k]++;
            if (D[k] > max_deg) max_deg = D[k];
            if (k != j) {
                int found = 0;
                for (l = IA[k]; l < IA[k + 1]; l++) {
                    if (JA[l] == j) { found = 1; break; }
                    else if (JA[l] > j) break;
                }
                if (!found) {
                    D[j]++;
                    if (D[j] > max_deg) max_deg = D[j];
                }
            }
        }
    if (max_deg == 0) {
        for (i = 0; i < N; i++) P[i] = i;
        free(c2); free(r2); free(D);
        return;
    }
    rcm_graph g = {IA, JA, c2, r2};
    rcm_rec *S = (rcm_rec *)calloc((size_t)max_deg + 1, sizeof(rcm_rec));
    rcm_rec *Q = (rcm_rec *)calloc((size_t)N + 1, sizeof(rcm_rec));
    rcm_rec *Q2 = (rcm_rec *)calloc((size_t)N + 1, sizeof(rcm_rec));
    int *visit = (int *)calloc((size_t)N, sizeof(int));
    int *visit2 = (int *)calloc((size_t)N, sizeof(int));
    int *nb = (int *)calloc((size_t)2 * max_deg + 2, sizeof(int));
    int c = -1, qh = 0, qt = 0, s = 0;
    /* Main loop (symrcm.c:71-220). The bandwidth bookkeeping there does not affect P. */
    do {
        for (i = 0; i < N; i++)
            if (!visit[i]) break;
        rcm_rec v;
        v.id = rcm_start(N, &g, D, i, Q2, visit2, nb);
        v.deg = D[v.id];
        v.dist = 0;
        visit[v.id] = 1;
        Q[qt] = v; qt = (qt + 1) % (N + 1);
        while (qh != qt) {
            v = Q[qh]; qh = (qh + 1) % (N + 1);
            P[++c] = v.id;
            s = 0;
            const int nn = rcm_neigh(&g, v.id, nb);
            for (k = 0; k < nn; k++) {
                const int r = nb[k];
                if (!visit[r]) {
                    rcm_rec ww = {r, D[r], v.dist + 1};
                    heap_push(S, &s, ww);
                    visit[r] = 1;
                }
            }
            while (s != 0) {
                rcm_rec u = heap_pop(S, &s);
                Q[qt] = u; qt = (qt + 1) % (N + 1);
            }
        }
    } while (c + 1 < N);
    free(S); free(Q); free(Q2); free(visit); free(visit2); free(nb);
    free(c2); free(r2); free(D);
    /* reverse (symrcm.c:231-235) */
    for (i = 0, j = N - 1; i <= N / 2 - 1; i++, j--) {
        const int t = P[i]; P[i] = P[j]; P[j] = t;
    }
}

/* ====================================================================== CSR permutation */

typedef struct { int slot, col, row; } csr_trip;

static int trip_cmp(const void *a, const void *b)
{
    const csr_trip *x = (const csr_trip *)a, *y = (const csr_trip *)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return 0;
}

void orc_csr_permute(int n, int nnz, int *IA, int *JA, const int *perm, int *pT)
{
    int i, j, k;
    /* MATRIX_ROW_permutation (reordering.c:139-167): new row i is old row perm[i]. */
    int *apT = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    int *aJA = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    int *aIA = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    aIA[0] = 0;
    for (i = 0, k = 0; i < n; i++) {
        for (j = IA[perm[i]]; j < IA[perm[i] + 1]; j++, k++) { apT[k] = pT[j]; aJA[k] = JA[j]; }
        aIA[i + 1] = k;
    }
    memcpy(pT, apT, sizeof(int) * (size_t)nnz);
    memcpy(JA, aJA, sizeof(int) * (size_t)nnz);
    memcpy(IA, aIA, sizeof(int) * ((size_t)n + 1));
    free(apT); free(aJA); free(aIA);
    /* MATRIX_COL_permutation (reordering.c:101-134): columns renamed through the inverse
     * permutation, rows re-sorted by column; (row, col) pairs are unique so the order is total. */
    int *q = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    for (i = 0; i < n; i++) q[perm[i]] = i;
    csr_trip *t = (csr_trip *)malloc(sizeof(csr_trip) * ((size_t)nnz + 1));
    for (i = 0, k = 0; i < n; i++)
        for (j = IA[i]; j < IA[i + 1]; j++, k++) { t[k].slot = pT[j]; t[k].col = q[JA[j]]; t[k].row = i; }
    qsort(t, (size_t)nnz, sizeof(csr_trip), trip_cmp);
    for (i = 0; i < nnz; i++) { pT[i] = t[i].slot; JA[i] = t[i].col; }
    free(t); free(q);
}

/* ====================================================================== CSR structure */

/* Sorted unique column lists of the local block: for every element, every local row lm[a] < neq
 * receives every column lm[b] < neq of the same element (csr_Initialization.c:36-95, generalised
 * loop of 04_SS_NavierStokes_Equations/2D/01_CommonFiles/csr_Initialization.c:36-56). Built as
 * per-row counting + sort + unique instead of linked-list inserts; same sets, same order. */
static int csr_pattern(int nel, int s, int neq, const int *lm, int *IA, int *JA)
{
    int e, a, b, i;
    int *cnt = (int *)calloc((size_t)neq + 1, sizeof(int));
    for (e = 0; e < nel; e++)
        for (a = 0; a < s; a++) {
            const int I = lm[e * s + a];
            if (I < 0 || I >= neq) continue;
            for (b = 0; b < s; b++) {
                const int J = lm[e * s + b];
                if (J >= 0 && J < neq) cnt[I]++;
            }
        }
    int *ptr = (int *)malloc(sizeof(int) * ((size_t)neq + 1));
    ptr[0] = 0;
    for (i = 0; i < neq; i++) ptr[i + 1] = ptr[i] + cnt[i];
    int *buf = (int *)malloc(sizeof(int) * ((size_t)ptr[neq] + 1));
    int *fill = (int *)calloc((size_t)neq + 1, sizeof(int));
    for (e = 0; e < nel; e++)
        for (a = 0; a < s; a++) {
            const int I = lm[e * s + a];
            if (I < 0 || I >= neq) continue;
            for (b = 0; b < s; b++) {
                const int J = lm[e * s + b];
                if (J >= 0 && J < neq) buf[ptr[I] + fill[I]++] = J;
            }
        }
    int nnz = 0;
    if (IA) IA[0] = 0;
    for (i = 0; i < neq; i++) {
        int *r = buf + ptr[i];
        const int len = cnt[i];
        /* insertion sort: rows hold a handful of entries */
        for (a = 1; a < len; a++) {
            const int v = r[a];
            b = a - 1;
            while (b >= 0 && r[b] > v) { r[b + 1] = r[b]; b--; }
            r[b + 1] = v;
        }
        for (a = 0; a < len; a++) {
            if (a > 0 && r[a] == r[a - 1]) continue;
            if (JA) JA[nnz] = r[a];
            nnz++;
        }
        if (IA) IA[i + 1] = nnz;
    }
    free(cnt); free(ptr); free(buf); free(fill);
    return nnz;
}

int orc_csr_nnz(int nel, int s, int neq, const int *lm)
{
    return csr_pattern(nel, s, neq, lm, NULL, NULL);
}

static int row_search(const int *IA, const int *JA, int row, int col)
{
    int lo = IA[row], hi = IA[row + 1] - 1;
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        if (JA[m] == col) return m;
        if (JA[m] < col) lo = m + 1; else hi = m - 1;
    }
    return -1;
}

int orc_csr_init(int nel, int s, int neq, int *lm, int reorder, int permute_lm,
                 int ndofs_total, const int *type, int *invP_id,
                 int *IA, int *JA, int *scheme, int *perm)
{
    int e, a, b, i;
    if (reorder != 0 && reorder != 1) return -1;
    const int nnz = csr_pattern(nel, s, neq, lm, IA, JA);
    /* Scheme_by_Element: csr_Initialization.c:101-191 (dummy slot nnz where a row or column is
     * Dirichlet). */
    for (e = 0; e < nel; e++)
        for (a = 0; a < s; a++) {
            const int I = lm[e * s + a];
            for (b = 0; b < s; b++) {
                const int J = lm[e * s + b];
                int slot = nnz;
                if (I >= 0 && I < neq && J >= 0 && J < neq) slot = row_search(IA, JA, I, J);
                scheme[(size_t)e * s * s + a * s + b] = slot;
            }
        }
    for (i = 0; i < neq; i++) perm[i] = i;
    /* reordering (reordering.c:3-22), then the permutations (csr_Initialization.c:421-433). */
    if (reorder == 1) orc_symrcm(neq, IA, JA, perm);
    int *pT = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    for (i = 0; i < nnz; i++) pT[i] = i;
    orc_csr_permute(neq, nnz, IA, JA, perm, pT);
    /* invPermCSR / invperm and the remaps (csr_Initialization.c:435-458). */
    int *ipT = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    int *iperm = (int *)malloc(sizeof(int) * ((size_t)neq + 1));
    for (i = 0; i < nnz; i++) ipT[pT[i]] = i;
    ipT[nnz] = nnz;
    for (i = 0; i < neq; i++) iperm[perm[i]] = i;
    iperm[neq] = neq;
    const size_t ss = (size_t)s * s;
    for (e = 0; e < nel; e++) {
        for (a = 0; a < (int)ss; a++) scheme[e * ss + a] = ipT[scheme[e * ss + a]];
        if (permute_lm)
            for (a = 0; a < s; a++)
                if (lm[e * s + a] >= 0 && lm[e * s + a] < neq) lm[e * s + a] = iperm[lm[e * s + a]];
    }
    if (invP_id)
        for (i = 0; i < ndofs_total; i++)
            if (type[i] == 1) invP_id[i] = iperm[invP_id[i]];
    free(pT); free(ipT); free(iperm);
    return nnz;
}

int orc_ebe_permutation(int nel, int NEQ, const int *ien, const int *type, const int *invP_id,
                        int reorder, int *Perm, int *invPerm)
{
    int i;
    /* Adjacency of unknowns (Permutation_of_LM.c:50-80): for each element every unknown vertex
     * gets every unknown vertex as a column; the same sorted unique lists as csr_pattern on the
     * element map of the unknowns. */
    int *lmu = (int *)malloc(sizeof(int) * (size_t)nel * 3 + 1);
    for (i = 0; i < 3 * nel; i++) {
        const int J = ien[i];
        lmu[i] = (type[J] != 0) ? invP_id[J] : -1;
    }
    const int nnz = csr_pattern(nel, 3, NEQ, lmu, NULL, NULL);
    int *IA = (int *)malloc(sizeof(int) * ((size_t)NEQ + 1));
    int *JA = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    csr_pattern(nel, 3, NEQ, lmu, IA, JA);
    for (i = 0; i < NEQ; i++) Perm[i] = i;
    Perm[NEQ] = NEQ;
    if (reorder == 1) orc_symrcm(NEQ, IA, JA, Perm);
    for (i = 0; i < NEQ; i++) invPerm[Perm[i]] = i;
    invPerm[NEQ] = NEQ;
    free(lmu); free(IA); free(JA);
    return nnz;
}

/* ====================================================================== transport assembly */

/* CAU_ShockCapture, steady form (01_.../CAU_ShockCapture.c:3-31; transient form adds dueb to
 * the residual, 02_.../CAU_ShockCapture.c:3-31). */
static double cau_eu(double kx, double ky, double Be_x, double Be_y, double gamma, double ue1,
                     double ue2, double ue3, double ueb, double dueb, double feb, double y23,
                     double y31, double y12, double x32, double x13, double x21, double invArea,
                     double h)
{
    const double third = 1. / 3.;
    double BetaGradu = 0.5 * invArea * (ue1 * (Be_x * y23 + Be_y * x32) + ue2 * (Be_x * y31 + Be_y * x13) +
                                        ue3 * (Be_x * y12 + Be_y * x21));
    double Ru = dueb + BetaGradu + gamma * ueb - feb;
    double g0 = 0.5 * invArea * (y23 * ue1 + y31 * ue2 + y12 * ue3);
    double g1 = 0.5 * invArea * (x32 * ue1 + x13 * ue2 + x21 * ue3);
    double normGradu = sqrt(g0 * g0 + g1 * g1);
    if (!(normGradu > 0)) return 0;
    double Be_xp = BetaGradu * g0 / (normGradu * normGradu);
    double Be_yp = BetaGradu * g1 / (normGradu * normGradu);
    double ktil_p = Be_xp * Be_xp * kx + Be_yp * Be_yp * ky;
    double normBe_p = sqrt(Be_xp * Be_xp + Be_yp * Be_yp);
    double Pee = 0.5 * h * normBe_p * normBe_p * normBe_p / ktil_p;
    if (third * Pee < 0.7)
        return 0.5 * h * third * Pee * fabs(Ru) / normGradu;
    return 0.5 * h * 0.7 * fabs(Ru) / normGradu;
}

/* Per-element SUPG block shared by the steady and the transient builds
 * (Build_K_F_SUPG.c:60-152, Build_M_K_R_SUPG.c:60-166). */
typedef struct {
    double X[3], Y[3], y23, y31, y12, x32, x13, x21, Area, invArea, tau, Be_x, Be_y, fe[3], feb;
    double ke[3][3];
} supg_elem;

static void supg_geometry(supg_elem *E, const orc_transport_problem *pb)
{
    const double third = 1.0 / 3.0;
    E->y23 = E->Y[1] - E->Y[2];
    E->y31 = E->Y[2] - E->Y[0];
    E->y12 = E->Y[0] - E->Y[1];
    E->x32 = E->X[2] - E->X[1];
    E->x13 = E->X[0] - E->X[2];
    E->x21 = E->X[1] - E->X[0];
    const double TwoA = E->x21 * E->y31 - E->x13 * E->y12;
    E->Area = 0.5 * TwoA;
    E->invArea = 1.0 / E->Area;
    const double h = sqrt(fabs(TwoA));
    const double XB = third * (E->X[0] + E->X[1] + E->X[2]);
    const double YB = third * (E->Y[0] + E->Y[1] + E->Y[2]);
    E->Be_x = pb->bx[0] + pb->bx[1] * XB + pb->bx[2] * YB;
    E->Be_y = pb->by[0] + pb->by[1] * XB + pb->by[2] * YB;
    const double norm = sqrt(E->Be_x * E->Be_x + E->Be_y * E->Be_y);
    E->tau = (norm > 0) ? 0.5 * h / norm : 0;
    E->fe[0] = E->fe[1] = E->fe[2] = pb->f0;
    E->feb = third * (E->fe[0] + E->fe[1] + E->fe[2]);
}

static void supg_ke(supg_elem *E, const orc_transport_problem *pb, double Eu)
{
    const double sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    const double kx = pb->k, ky = pb->k, gamma = pb->gamma;
    const double Be_x = E->Be_x, Be_y = E->Be_y, tau = E->tau, invArea = E->invArea, Area = E->Area;
    const double y23 = E->y23, y31 = E->y31, y12 = E->y12, x32 = E->x32, x13 = E->x13, x21 = E->x21;
    double D12 = 0.25 * invArea * ((kx * y23 * y31) + (ky * x32 * x13));
    double D13 = 0.25 * invArea * ((kx * y23 * y12) + (ky * x32 * x21));
    double D23 = 0.25 * invArea * ((kx * y31 * y12) + (ky * x13 * x21));
    double C11 = sixth * (Be_x * y23 + Be_y * x32);
    double C12 = sixth * (Be_x * y31 + Be_y * x13);
    double C13 = sixth * (Be_x * y12 + Be_y * x21);
    double R11 = sixth * (gamma * Area);
    double R12 = twelfth * (gamma * Area);
    double Cpg12 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y31 + Be_x * Be_y * x13) + x32 * (Be_x * Be_y * y31 + Be_y * Be_y * x13));
    double Cpg13 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) + x32 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
    double Cpg23 = 0.25 * invArea * tau * (y31 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) + x13 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
    double Rpg11 = sixth * tau * gamma * (Be_x * y23 + Be_y * x32);
    double Rpg21 = sixth * tau * gamma * (Be_x * y31 + Be_y * x13);
    double Rpg31 = sixth * tau * gamma * (Be_x * y12 + Be_y * x21);
    double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
    double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
    double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);
    E->ke[0][0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
    E->ke[0][1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
    E->ke[0][2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
    E->ke[1][0] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
    E->ke[1][1] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
    E->ke[1][2] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
    E->ke[2][0] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
    E->ke[2][1] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
    E->ke[2][2] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);
}

void orc_build_transport_supg(int nnodes, int nel, int neq, const double *x, const double *y,
                              const int *type, const int *invP_id, const int *ien,
                              const double *ubc, const double *u, const orc_transport_problem *pb,
                              double *Ke, double *F)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    int J, e, a;
    for (J = 0; J < neq; J++) F[J] = 0;
    double *U = (double *)malloc(sizeof(double) * ((size_t)nnodes + 1));
    for (J = 0; J < nnodes; J++) U[J] = (type[J] == 0) ? ubc[J] : u[invP_id[J]];
    for (e = 0; e < nel; e++) {
        supg_elem E;
        const int *v = ien + 3 * e;
        for (a = 0; a < 3; a++) { E.X[a] = x[v[a]]; E.Y[a] = y[v[a]]; }
        supg_geometry(&E, pb);
        const double ue1 = U[v[0]], ue2 = U[v[1]], ue3 = U[v[2]];
        const double ueb = third * (ue1 + ue2 + ue3);
        double Eu = 0;
        if (pb->shock == 1) {
            const double hs = sqrt(fabs(2 * E.Area));
            Eu = cau_eu(pb->k, pb->k, E.Be_x, E.Be_y, pb->gamma, ue1, ue2, ue3, ueb, 0.0, E.feb,
                        E.y23, E.y31, E.y12, E.x32, E.x13, E.x21, E.invArea, hs);
        }
        supg_ke(&E, pb, Eu);
        const double fs = E.fe[0] + E.fe[1] + E.fe[2];
        double Fe[3];
        Fe[0] = E.Area * twelfth * (2 * E.fe[0] + E.fe[1] + E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y23 + E.Be_y * E.x32);
        Fe[1] = E.Area * twelfth * (E.fe[0] + 2 * E.fe[1] + E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y31 + E.Be_y * E.x13);
        Fe[2] = E.Area * twelfth * (E.fe[0] + E.fe[1] + 2 * E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y12 + E.Be_y * E.x21);
        /* F_assembly.c:5-27: own load, then Dirichlet lifting in vertex order. */
        for (a = 0; a < 3; a++) {
            if (type[v[a]] != 1) continue;
            const int I = invP_id[v[a]];
            F[I] += Fe[a];
            for (int b = 0; b < 3; b++)
                if (b != a && type[v[b]] == 0) F[I] -= E.ke[a][b] * ubc[v[b]];
        }
        for (a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) Ke[(size_t)9 * e + 3 * a + b] = E.ke[a][b];
    }
    free(U);
}

void orc_build_transport_transient(int nnodes, int nel, int NEQ, const double *x, const double *y,
                                   const int *type, const int *invP_id, const int *ien,
                                   const int *lm, const double *ubc, const double *u,
                                   const double *du, const orc_transport_problem *pb,
                                   double alpha, double dt, double *Ae, double *R)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    int J, e, a, b;
    for (J = 0; J <= NEQ; J++) R[J] = 0;
    double *U = (double *)malloc(sizeof(double) * ((size_t)nnodes + 1));
    double *dU = (double *)malloc(sizeof(double) * ((size_t)nnodes + 1));
    for (J = 0; J < nnodes; J++) {
        if (type[J] == 0) { U[J] = ubc[J]; dU[J] = 0.0; }
        else { U[J] = u[invP_id[J]]; dU[J] = du[invP_id[J]]; }
    }
    for (e = 0; e < nel; e++) {
        supg_elem E;
        const int *v = ien + 3 * e;
        for (a = 0; a < 3; a++) { E.X[a] = x[v[a]]; E.Y[a] = y[v[a]]; }
        supg_geometry(&E, pb);
        double ue[3] = {U[v[0]], U[v[1]], U[v[2]]}, due[3] = {dU[v[0]], dU[v[1]], dU[v[2]]};
        const double ueb = third * (ue[0] + ue[1] + ue[2]);
        const double dueb = third * (due[0] + due[1] + due[2]);
        double Eu = 0;
        if (pb->shock == 1) {
            const double hs = sqrt(fabs(2 * E.Area));
            Eu = cau_eu(pb->k, pb->k, E.Be_x, E.Be_y, pb->gamma, ue[0], ue[1], ue[2], ueb, dueb,
                        E.feb, E.y23, E.y31, E.y12, E.x32, E.x13, E.x21, E.invArea, hs);
        }
        supg_ke(&E, pb, Eu);
        const double sixA = sixth * E.Area, twA = twelfth * E.Area;
        const double C11 = sixth * (E.Be_x * E.y23 + E.Be_y * E.x32);
        const double C12 = sixth * (E.Be_x * E.y31 + E.Be_y * E.x13);
        const double C13 = sixth * (E.Be_x * E.y12 + E.Be_y * E.x21);
        double me[3][3];
        /* Build_M_K_R_SUPG.c:148-156 */
        me[0][0] = sixA + E.tau * C11; me[0][1] = twA + E.tau * C11; me[0][2] = twA + E.tau * C11;
        me[1][0] = twA + E.tau * C12;  me[1][1] = sixA + E.tau * C12; me[1][2] = twA + E.tau * C12;
        me[2][0] = twA + E.tau * C13;  me[2][1] = twA + E.tau * C13;  me[2][2] = sixA + E.tau * C13;
        for (a = 0; a < 3; a++)
            for (b = 0; b < 3; b++) Ae[(size_t)9 * e + 3 * a + b] = me[a][b] + alpha * dt * E.ke[a][b];
        const double fs = E.fe[0] + E.fe[1] + E.fe[2];
        const int *L = lm + 3 * e;
        /* :173-184 */
        R[L[0]] += E.Area * twelfth * (2 * E.fe[0] + E.fe[1] + E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y23 + E.Be_y * E.x32);
        R[L[1]] += E.Area * twelfth * (E.fe[0] + 2 * E.fe[1] + E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y31 + E.Be_y * E.x13);
        R[L[2]] += E.Area * twelfth * (E.fe[0] + E.fe[1] + 2 * E.fe[2]) + sixth * E.tau * fs * (E.Be_x * E.y12 + E.Be_y * E.x21);
        for (a = 0; a < 3; a++) {
            const int I = L[a];
            for (b = 0; b < 3; b++) R[I] -= me[a][b] * due[b] + E.ke[a][b] * ue[b];
        }
        R[NEQ] = 0;
    }
    free(U); free(dU);
}

/* ====================================================================== Navier-Stokes (NDOF = 3) */

void orc_ns_nodal_state(int nnodes, const int *type3, const int *invP3, const double *u,
                        const double *presc3, double *U3)
{
    for (int k = 0; k < 3 * nnodes; k++) U3[k] = (type3[k] > 0) ? u[invP3[k]] : presc3[k];
}

/* Per-element restatement of Build_K_F_SUPG_PSPG.c:41-466. Vertex a has the edge vectors
 * b[a] = (y23, y31, y12)[a] and c[a] = (x32, x13, x21)[a]; every term below is the reference's
 * expression for vertex pair (a, b) evaluated in the reference's operation order (for the
 * symmetric terms the reference evaluates the (min, max) pair and reuses it, so the loops do too). */
void orc_build_ns_supg_pspg(int nnodes, int nel, int NEQ, const double *x, const double *y,
                            const int *ien, const int *lm, const double *U3, const double *fx,
                            const double *fy, double Re, double *Ke, double *F)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    (void)nnodes;
    for (int I = 0; I <= NEQ; I++) F[I] = 0;
    for (int e = 0; e < nel; e++) {
        const int *v = ien + 3 * e;
        const double X0 = x[v[0]], X1 = x[v[1]], X2 = x[v[2]];
        const double Y0 = y[v[0]], Y1 = y[v[1]], Y2 = y[v[2]];
        const double b[3] = {Y1 - Y2, Y2 - Y0, Y0 - Y1};   /* y23, y31, y12 */
        const double c[3] = {X2 - X1, X0 - X2, X1 - X0};   /* x32, x13, x21 */
        const double TwoA = fabs(c[2] * b[1] - c[1] * b[2]);
        const double Area = 0.5 * TwoA, invArea = 1.0 / Area;
        double ux_[3], uy_[3], p_[3];
        for (int a = 0; a < 3; a++) {
            ux_[a] = U3[3 * v[a]];
            uy_[a] = U3[3 * v[a] + 1];
            p_[a] = U3[3 * v[a] + 2];
        }
        const double ux = third * (ux_[0] + ux_[1] + ux_[2]);
        const double uy = third * (uy_[0] + uy_[1] + uy_[2]);
        const double duxdx = 0.5 * invArea * (ux_[0] * b[0] + ux_[1] * b[1] + ux_[2] * b[2]);
        const double duxdy = 0.5 * invArea * (ux_[0] * c[0] + ux_[1] * c[1] + ux_[2] * c[2]);
        const double duydx = 0.5 * invArea * (uy_[0] * b[0] + uy_[1] * b[1] + uy_[2] * b[2]);
        const double duydy = 0.5 * invArea * (uy_[0] * c[0] + uy_[1] * c[1] + uy_[2] * c[2]);
        const double duxduy = ux_[0] * b[0] + uy_[0] * c[0] + ux_[1] * b[1] + uy_[1] * c[1] + ux_[2] * b[2] +
                              uy_[2] * c[2];
        const double p = third * (p_[0] + p_[1] + p_[2]);
        const double dpdx = 0.5 * invArea * (p_[0] * b[0] + p_[1] * b[1] + p_[2] * b[2]);
        const double dpdy = 0.5 * invArea * (p_[0] * c[0] + p_[1] * c[1] + p_[2] * c[2]);
        double C[3];
        for (int a = 0; a < 3; a++) C[a] = 0.5 * invArea * (ux * b[a] + uy * c[a]);
        /* stabilisation parameters (:118-131) */
        const double h = sqrt(TwoA), visc = 1. / Re, rho = 1.;
        const double invrho = 1.0 / rho, nu = visc / rho;
        const double unorm = sqrt(ux * ux + uy * uy);
        double auxtau = (2.0 * unorm / h) * (2.0 * unorm / h) + 9.0 * (4.0 * nu / (h * h)) * (4.0 * nu / (h * h));
        auxtau = sqrt(auxtau);
        const double tau = 1.0 / auxtau;
        const double tau_l = sqrt(nu * nu + (0.5 * h * unorm) * (0.5 * h * unorm));
        /* dof-interleaved edge vector P = (y23, x32, y31, x13, y12, x21) of the LSIC terms */
        const double P6[6] = {b[0], c[0], b[1], c[1], b[2], c[2]};
        double *K = Ke + (size_t)81 * e;
        for (int a = 0; a < 3; a++) {
            for (int bb = 0; bb < 3; bb++) {
                const int lo = a < bb ? a : bb, hi = a < bb ? bb : a;
                const double Nb = rho * third * Area * C[bb];                       /* N11/N13/N15 */
                const double Nd = rho * tau * Area * C[lo] * C[hi];                 /* Ndelta */
                const double Kuu = visc * 0.25 * invArea * (2 * b[lo] * b[hi] + c[lo] * c[hi]);
                const double Kvv = visc * 0.25 * invArea * (b[lo] * b[hi] + 2 * c[lo] * c[hi]);
                const double Kuv = visc * 0.25 * invArea * (c[a] * b[bb]);
                const double Kvu = visc * 0.25 * invArea * (b[a] * c[bb]);
                const double ks = tau_l * rho * 0.25 * invArea;
                const double Ks_uu = ks * P6[2 * bb] * P6[2 * a], Ks_uv = ks * P6[2 * bb + 1] * P6[2 * a];
                const double Ks_vu = ks * P6[2 * bb] * P6[2 * a + 1], Ks_vv = ks * P6[2 * bb + 1] * P6[2 * a + 1];
                const double GTu_a = sixth * b[a], GTv_a = sixth * c[a];
                const double GTu_b = sixth * b[bb], GTv_b = sixth * c[bb];
                const double Gdu_ab = 0.5 * tau * C[a] * b[bb], Gdv_ab = 0.5 * tau * C[a] * c[bb];
                const double Gdu_ba = 0.5 * tau * C[bb] * b[a], Gdv_ba = 0.5 * tau * C[bb] * c[a];
                const double Gphi = invrho * 0.25 * tau * invArea * (b[lo] * b[hi] + c[lo] * c[hi]);
                double *r0 = K + 27 * a + 3 * bb;   /* row 3a (u), columns 3bb .. 3bb+2 */
                r0[0] = Nb + Nd + Kuu + Ks_uu;
                r0[1] = Kuv + Ks_uv;
                r0[2] = -(GTu_a + Gdu_ab);
                double *r1 = r0 + 9;                /* row 3a+1 (v) */
                r1[0] = Kvu + Ks_vu;
                r1[1] = Nb + Nd + Kvv + Ks_vv;
                r1[2] = -(GTv_a + Gdv_ab);
                double *r2 = r0 + 18;               /* row 3a+2 (p) */
                r2[0] = GTu_b + Gdu_ba;
                r2[1] = GTv_b + Gdv_ba;
                r2[2] = Gphi;
            }
        }
        /* loads (:392-430) */
        const double fxB = third * (fx[v[0]] + fx[v[1]] + fx[v[2]]);
        const double fyB = third * (fy[v[0]] + fy[v[1]] + fy[v[2]]);
        const double f1 = rho * third * Area * fxB, f2 = rho * third * Area * fyB;
        /* residual (:432-464) */
        const double cvx = duxdx * ux + duxdy * uy, cvy = duydx * ux + duydy * uy;
        const double Nvx = rho * third * Area * cvx, Nvy = rho * third * Area * cvy;
        const double GTv = third * Area * (duxdx + duydy);
        double Res[9];
        for (int a = 0; a < 3; a++) {
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
            Res[3 * a] = (f1 + fdx) - (Nvx + Ndx + Kvx + Ksx - (Gvx + Gdx));
            Res[3 * a + 1] = (f2 + fdy) - (Nvy + Ndy + Kvy + Ksy - (Gvy + Gdy));
            Res[3 * a + 2] = fphi - (GTv + Nphi + Gphiv);
        }
        for (int i = 0; i < 9; i++) F[lm[9 * e + i]] += Res[i];
        F[NEQ] = 0;
    }
}

/* ====================================================================== Euler (NDOF = 4) */

void orc_euler_nodal_state(int nnodes, const int *type4, const int *invP4, const double *u,
                           const double *du, const double *presc4, double *U4, double *dU4)
{
    for (int k = 0; k < 4 * nnodes; k++) {
        if (type4[k] > 0) { U4[k] = u[invP4[k]]; dU4[k] = du[invP4[k]]; }
        else { U4[k] = presc4[k]; dU4[k] = 0.0; }
    }
}

/* Flux Jacobians at the barycentric state (Ax_Ay_calculations.c:3-53 dimensionless, :55-99
 * dimensional). Entries the reference leaves structurally zero are exact zeros here, including
 * Ay[0][3], which the reference's hand-expanded products treat as zero. */
static void euler_jacobians(const orc_euler_params *p, const double Ub[4], double Ax[4][4], double Ay[4][4])
{
    const double g = p->gamma;
    const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
    memset(Ax, 0, sizeof(double) * 16);
    memset(Ay, 0, sizeof(double) * 16);
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
    if (p->dimensionless) {
        const double nv = v1 * v1 + v2 * v2, eps2 = p->Mach * p->Mach;
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

/* C = A B, each entry accumulated over k = 0..3 from the k = 0 term. The reference's hand-expanded
 * products (Build_M_K_F_SUPG.c:163-240) sum the same terms in the same order without the terms
 * with a structural zero factor (and with factors 1 written out), which leaves every value equal. */
static void mul4(const double A[4][4], const double B[4][4], double Cm[4][4])
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = A[i][0] * B[0][j];
            for (int k = 1; k < 4; k++) acc += A[i][k] * B[k][j];
            Cm[i][j] = acc;
        }
}

/* w^T A0 w in the reference's grouping (Delta_CAU.c:39-73). */
static double a0_form(const double A0[4][4], const double *w)
{
    return w[0] * (A0[0][0] * w[0] + 2.0 * A0[0][1] * w[1] + 2.0 * A0[0][2] * w[2] + 2.0 * A0[0][3] * w[3]) +
           w[1] * (A0[1][1] * w[1] + 2.0 * A0[1][2] * w[2] + 2.0 * A0[1][3] * w[3]) +
           w[2] * (A0[2][2] * w[2] + 2.0 * A0[2][3] * w[3]) + w[3] * A0[3][3] * w[3];
}

/* Delta_CAU (03_Euler_Equations/2D/01_CommonFiles/Delta_CAU.c:4-83). */
static double euler_cau(double tolerance, const double gUx[4], const double gUy[4], const double Ax[4][4],
                        const double Ay[4][4], const double A0[4][4], const double dUb[4], const double b[3],
                        const double c[3], double twoArea)
{
    double AgUx[4], AgUy[4], gkx[4], gky[4], LUh[4];
    AgUx[0] = gUx[1];
    AgUx[1] = Ax[1][0] * gUx[0] + Ax[1][1] * gUx[1] + Ax[1][2] * gUx[2] + Ax[1][3] * gUx[3];
    AgUx[2] = Ax[2][0] * gUx[0] + Ax[2][1] * gUx[1] + Ax[2][2] * gUx[2];
    AgUx[3] = Ax[3][0] * gUx[0] + Ax[3][1] * gUx[1] + Ax[3][2] * gUx[2] + Ax[3][3] * gUx[3];
    AgUy[0] = gUy[2];
    AgUy[1] = Ay[1][0] * gUy[0] + Ay[1][1] * gUy[1] + Ay[1][2] * gUy[2];
    AgUy[2] = Ay[2][0] * gUy[0] + Ay[2][1] * gUy[1] + Ay[2][2] * gUy[2] + Ay[2][3] * gUy[3];
    AgUy[3] = Ay[3][0] * gUy[0] + Ay[3][1] * gUy[1] + Ay[3][2] * gUy[2] + Ay[3][3] * gUy[3];
    for (int i = 0; i < 4; i++) {
        gkx[i] = (b[0] * gUx[i] + c[0] * gUy[i]) / twoArea;
        gky[i] = (b[1] * gUx[i] + c[1] * gUy[i]) / twoArea;
    }
    const double n_gk = sqrt(a0_form(A0, gkx) + a0_form(A0, gky));
    for (int i = 0; i < 4; i++) LUh[i] = dUb[i] + AgUx[i] + AgUy[i];
    const double n_LUh = sqrt(a0_form(A0, LUh));
    const double n_gU = sqrt(a0_form(A0, gUx) + a0_form(A0, gUy));
    return (n_gU >= tolerance) ? n_LUh / n_gk : 0.0;
}

void orc_build_euler_supg(int nnodes, int nel, int NEQ, const double *x, const double *y,
                          const int *ien, const int *lm, const double *U4, const double *dU4,
                          const orc_euler_params *p, double *Ae, double *R)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    const double gamma = p->gamma, alpha = p->alpha, delta_t = p->dt;
    (void)nnodes;
    for (int I = 0; I <= NEQ; I++) R[I] = 0;
    for (int e = 0; e < nel; e++) {
        const int *v = ien + 3 * e;
        const double X0 = x[v[0]], X1 = x[v[1]], X2 = x[v[2]];
        const double Y0 = y[v[0]], Y1 = y[v[1]], Y2 = y[v[2]];
        const double b[3] = {Y1 - Y2, Y2 - Y0, Y0 - Y1};   /* y23, y31, y12 */
        const double c[3] = {X2 - X1, X0 - X2, X1 - X0};   /* x32, x13, x21 */
        const double twoArea = fabs(c[2] * b[1] - c[1] * b[2]);
        const double Area = 0.5 * twoArea;
        double Ue[12], dUe[12], Ub[4], dUb[4], gUx[4], gUy[4];
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 4; k++) {
                Ue[4 * a + k] = U4[4 * v[a] + k];
                dUe[4 * a + k] = dU4[4 * v[a] + k];
            }
        for (int i = 0; i < 4; i++) {
            Ub[i] = (Ue[i] + Ue[i + 4] + Ue[i + 8]) * third;
            dUb[i] = (dUe[i] + dUe[i + 4] + dUe[i + 8]) * third;
        }
        for (int i = 0; i < 4; i++) {
            gUx[i] = (Ue[i] * b[0] + Ue[i + 4] * b[1] + Ue[i + 8] * b[2]) / twoArea;
            gUy[i] = (Ue[i] * c[0] + Ue[i + 4] * c[1] + Ue[i + 8] * c[2]) / twoArea;
        }
        const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
        const double nU = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        double Ax[4][4], Ay[4][4];
        euler_jacobians(p, Ub, Ax, Ay);
        /* inverse Riemannian metric A0^-1 in entropy variables (:134-160) */
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
        A0[1][0] = A0[0][1]; A0[2][0] = A0[0][2]; A0[2][1] = A0[1][2];
        A0[3][0] = A0[0][3]; A0[3][1] = A0[1][3]; A0[3][2] = A0[2][3];
        double Axx[4][4], Axy[4][4], Ayx[4][4], Ayy[4][4];
        mul4(Ax, Ax, Axx);
        mul4(Ax, Ay, Axy);
        mul4(Ay, Ax, Ayx);
        mul4(Ay, Ay, Ayy);
        const double delta = p->shock ? euler_cau(p->tolerance, gUx, gUy, Ax, Ay, A0, dUb, b, c, twoArea) : 0.0;
        /* SUPG parameter (:246-271) */
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
        /* per-vertex blocks: C_a = (b_a Ax + c_a Ay)/6, its SUPG products, mass/convection (:276-720) */
        double Cv[3][4][4], AX[3][4][4], AY[3][4][4];
        for (int a = 0; a < 3; a++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    Cv[a][i][j] = (b[a] * Ax[i][j] + c[a] * Ay[i][j]) * sixth;
                    AX[a][i][j] = b[a] * Axx[i][j] + c[a] * Ayx[i][j];
                    AY[a][i][j] = b[a] * Axy[i][j] + c[a] * Ayy[i][j];
                }
        const double Mg1 = Area * sixth, Mg2 = Area / 12.0;
        const double tau4area = tau / (4 * Area), delta4area = delta / (4.0 * Area);
        double Kcau[3][3];
        for (int a = 0; a < 3; a++)
            for (int bb = a + 1; bb < 3; bb++)
                Kcau[a][bb] = Kcau[bb][a] = delta4area * (b[a] * b[bb] + c[a] * c[bb]);
        double Me[12][12], Ke[12][12];
        for (int a = 0; a < 3; a++) {
            const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;   /* the other vertices, ascending */
            for (int bb = 0; bb < 3; bb++)
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) {
                        const double Mpg = Cv[a][i][j] * tau;
                        double m, k;
                        if (i == j) m = (a == bb ? Mg1 : Mg2) + Mpg;
                        else m = Mpg;
                        if (a == bb) {
                            const double k1v = tau4area * (b[o1] * AX[a][i][j] + c[o1] * AY[a][i][j]);
                            const double k2v = tau4area * (b[o2] * AX[a][i][j] + c[o2] * AY[a][i][j]);
                            k = Cv[a][i][j] - (k1v + k2v);
                            if (i == j) k = k - (Kcau[a][o1] + Kcau[a][o2]);
                        } else {
                            const double kp = tau4area * (b[bb] * AX[a][i][j] + c[bb] * AY[a][i][j]);
                            k = Cv[bb][i][j] + kp;
                            if (i == j) k = k + Kcau[a][bb];
                        }
                        Me[4 * a + i][4 * bb + j] = m;
                        Ke[4 * a + i][4 * bb + j] = k;
                    }
        }
        /* Ae and the residual (:1025-1047) */
        double *A = Ae + (size_t)144 * e;
        for (int i = 0; i < 12; i++) {
            double MdU = 0, KU = 0;
            for (int j = 0; j < 12; j++) {
                MdU += Me[i][j] * dUe[j];
                KU += Ke[i][j] * Ue[j];
                A[12 * i + j] = Me[i][j] + alpha * delta_t * Ke[i][j];
            }
            R[lm[12 * e + i]] += -MdU - KU;
        }
        R[NEQ] = 0;
    }
}

void orc_csr_assemble(int nel, int s, int nnz, const int *scheme, const double *Ke, double *AA)
{
    const size_t ss = (size_t)s * s;
    int e;
    size_t a;
    for (a = 0; a <= (size_t)nnz; a++) AA[a] = 0;
    for (e = 0; e < nel; e++) {
        for (a = 0; a < ss; a++) AA[scheme[e * ss + a]] += Ke[e * ss + a];
        AA[nnz] = 0;
    }
}

/* ====================================================================== operators */

void orc_ebemv(int nel, int NEQ, const int *lm, const int *Perm, const int *invPerm,
               const double *A, const double *P, double *Q)
{
    int I;
    double *Pa = (double *)calloc((size_t)NEQ + 1, sizeof(double));
    double *Qa = (double *)calloc((size_t)NEQ + 1, sizeof(double));
    for (I = 0; I < NEQ; I++) Pa[I] = P[Perm[I]];
    for (I = 0; I < nel; I++) {
        const int i1 = invPerm[lm[3 * I]], i2 = invPerm[lm[3 * I + 1]], i3 = invPerm[lm[3 * I + 2]];
        const double p1 = Pa[i1], p2 = Pa[i2], p3 = Pa[i3];
        const double *a = A + (size_t)9 * I;
        Qa[i1] += a[0] * p1 + a[1] * p2 + a[2] * p3;
        Qa[i2] += a[3] * p1 + a[4] * p2 + a[5] * p3;
        Qa[i3] += a[6] * p1 + a[7] * p2 + a[8] * p3;
        Qa[NEQ] = 0;
    }
    for (I = 0; I < NEQ; I++) Q[Perm[I]] = Qa[I];
    free(Pa); free(Qa);
}

void orc_ebemv_block(int nel, int s, int NEQ, const int *lm, const double *A, const double *P,
                     double *Q)
{
    int e, a, b;
    for (a = 0; a < NEQ; a++) Q[a] = 0;
    double *p = (double *)malloc(sizeof(double) * (size_t)s);
    for (e = 0; e < nel; e++) {
        const int *L = lm + (size_t)e * s;
        const double *Ae = A + (size_t)e * s * s;
        for (b = 0; b < s; b++) p[b] = (L[b] < NEQ) ? P[L[b]] : 0.0;
        for (a = 0; a < s; a++) {
            if (L[a] >= NEQ) continue;
            double acc = 0;
            for (b = 0; b < s; b++) acc += Ae[a * s + b] * p[b];
            Q[L[a]] += acc;
        }
    }
    free(p);
}

void orc_csrmv(int n, const int *IA, const int *JA, const double *AA, const double *P, double *Q)
{
    int i, j;
    for (i = 0; i < n; i++) {
        double q = 0;
        for (j = IA[i]; j < IA[i + 1]; j++) q += AA[j] * P[JA[j]];
        Q[i] = q;
    }
}

/* ====================================================================== BLAS */

double orc_ddot(int n, const double *x, const double *y)
{
    long i, m = n - 4;
    double s = 0.0;
    for (i = 0; i < m; i += 5)
        s += x[i] * y[i] + x[i + 1] * y[i + 1] + x[i + 2] * y[i + 2] + x[i + 3] * y[i + 3] + x[i + 4] * y[i + 4];
    for (; i < n; i++) s += x[i] * y[i];
    return s;
}

void orc_daxpy(int n, double a, const double *x, double *y)
{
    int i;
    for (i = 0; i < n; i++) y[i] += a * x[i];
}

void orc_dscal(int n, double a, double *x)
{
    int i;
    for (i = 0; i < n; i++) x[i] = a * x[i];
}

/* ====================================================================== ILU(0) */

int orc_ilu0_factor(int n, const int *IA, const int *JA, const double *AA, double *LU, double *Dinv)
{
    int i, j, k;
    int *jw = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *ustart = (int *)malloc(sizeof(int) * ((size_t)n + 1)); /* first entry with column > row */
    int *lend = (int *)malloc(sizeof(int) * ((size_t)n + 1));   /* first entry with column >= row */
    for (i = 0; i < n; i++) {
        jw[i] = -1;
        lend[i] = IA[i + 1];
        ustart[i] = IA[i + 1];
        for (j = IA[i + 1] - 1; j >= IA[i]; j--) {
            if (JA[j] >= i) lend[i] = j;
            if (JA[j] > i) ustart[i] = j;
        }
        for (j = IA[i]; j < IA[i + 1]; j++) LU[j] = AA[j];
    }
    for (i = 0; i < n; i++) {
        /* ILUp (ILUp_precond_setup.c:84-166): row i in a work map, eliminate with the L entries
         * in ascending column order, multiplier scaled by the stored inverse pivot. The pivot
         * slot is always live (jw[i] = i there), even when A has no stored diagonal. */
        for (j = IA[i]; j < IA[i + 1]; j++) jw[JA[j]] = j;
        double d = (lend[i] < IA[i + 1] && JA[lend[i]] == i) ? LU[lend[i]] : 0.0;
        for (j = IA[i]; j < lend[i]; j++) {
            const int jrow = JA[j];
            LU[j] *= Dinv[jrow];
            const double lij = LU[j];
            for (k = ustart[jrow]; k < IA[jrow + 1]; k++) {
                const int col = JA[k];
                if (col == i) { d -= lij * LU[k]; continue; }
                const int pos = jw[col];
                if (pos == -1) continue;
                LU[pos] -= lij * LU[k];
            }
        }
        for (j = IA[i]; j < IA[i + 1]; j++) jw[JA[j]] = -1;
        if (d == 0) { free(jw); free(ustart); free(lend); return 1 + i; }
        if (lend[i] < IA[i + 1] && JA[lend[i]] == i) LU[lend[i]] = d;
        Dinv[i] = 1.0 / d;
    }
    free(jw); free(ustart); free(lend);
    return 0;
}

void orc_ilu0_apply(int n, const int *IA, const int *JA, const double *LU, const double *Dinv,
                    const double *p, double *z)
{
    int i, j;
    for (i = 0; i < n; i++) {
        double xi = p[i];
        for (j = IA[i]; j < IA[i + 1] && JA[j] < i; j++) xi -= z[JA[j]] * LU[j];
        z[i] = xi;
    }
    for (i = n - 1; i >= 0; i--) {
        double xi = z[i];
        for (j = IA[i]; j < IA[i + 1]; j++)
            if (JA[j] > i) xi -= z[JA[j]] * LU[j];
        z[i] = xi * Dinv[i];
    }
}

/* ====================================================================== nodal block diagonal */

/* 4x4 and 3x3 inverses by cofactors: matrix_Inverse (EBE/BlockDiag_precond_EBE_setup.c:92-147) and
 * DOF3_matrix_Inverse (EBE/BlockDiagDOF3_precond_EBE_setup.c), same expressions, same order. */
static void inv4(const double *B, double *R)
{
    const double a = B[0], b = B[1], c = B[2], d = B[3], e = B[4], f = B[5], g = B[6], h = B[7], i = B[8],
                 j = B[9], k = B[10], l = B[11], m = B[12], n = B[13], o = B[14], p = B[15];
    const double det = (a*f*k*p - a*f*l*o - a*g*j*p + a*g*l*n + a*h*j*o - a*h*k*n
                        - b*e*k*p + b*e*l*o + b*g*i*p - b*g*l*m - b*h*i*o + b*h*k*m +
                        c*(e*j*p - e*l*n - f*i*p + f*l*m + h*i*n - h*j*m) +
                        d*(-e*j*o + e*k*n + f*i*o - f*k*m - g*i*n + g*j*m));
    const double q = 1.0 / det;
    R[0] = q*(-h*k*n + g*l*n + h*j*o - f*l*o - g*j*p + f*k*p);
    R[1] = q*(d*k*n - c*l*n - d*j*o + b*l*o + c*j*p - b*k*p);
    R[2] = q*(-d*g*n + c*h*n + d*f*o - b*h*o - c*f*p + b*g*p);
    R[3] = q*(d*g*j - c*h*j - d*f*k + b*h*k + c*f*l - b*g*l);
    R[4] = q*(h*k*m - g*l*m - h*i*o + e*l*o + g*i*p - e*k*p);
    R[5] = q*(-d*k*m + c*l*m + d*i*o - a*l*o - c*i*p + a*k*p);
    R[6] = q*(d*g*m - c*h*m - d*e*o + a*h*o + c*e*p - a*g*p);
    R[7] = q*(-d*g*i + c*h*i + d*e*k - a*h*k - c*e*l + a*g*l);
    R[8] = q*(-h*j*m + f*l*m + h*i*n - e*l*n - f*i*p + e*j*p);
    R[9] = q*(d*j*m - b*l*m - d*i*n + a*l*n + b*i*p - a*j*p);
    R[10] = q*(-d*f*m + b*h*m + d*e*n - a*h*n - b*e*p + a*f*p);
    R[11] = q*(d*f*i - b*h*i - d*e*j + a*h*j + b*e*l - a*f*l);
    R[12] = q*(g*j*m - f*k*m - g*i*n + e*k*n + f*i*o - e*j*o);
    R[13] = q*(-c*j*m + b*k*m + c*i*n - a*k*n - b*i*o + a*j*o);
    R[14] = q*(c*j*m - b*g*m - c*e*n + a*g*n + b*e*o - a*f*o);
    R[15] = q*(-c*f*i + b*g*i + c*e*j - a*g*j - b*e*k + a*f*k);
}

static void inv3(const double *B, double *R)
{
    const double a = B[0], b = B[1], c = B[2], d = B[3], e = B[4], f = B[5], g = B[6], h = B[7], i = B[8];
    const double det = a*e*i - a*f*h - b*d*i + b*f*g + c*d*h - c*e*g;
    const double q = 1.0 / det;
    R[0] = q*(e*i - f*h); R[1] = q*(c*h - b*i); R[2] = q*(b*f - c*e);
    R[3] = q*(f*g - d*i); R[4] = q*(a*i - c*g); R[5] = q*(c*d - a*f);
    R[6] = q*(d*h - e*g); R[7] = q*(b*g - a*h); R[8] = q*(a*e - b*d);
}

void orc_blockdiag_setup(int nnodes, int nel, int ndof, const int *ien, const double *A, double *inv)
{
    const int s = 3 * ndof, bb = ndof * ndof;
    double *B = (double *)calloc((size_t)nnodes * bb, sizeof(double));
    /* BlockDiag_precond_EBE_setup.c:66-77: the diagonal block of vertex a of every element, added
     * to its node's block in element order */
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 3; a++) {
            const int v = ien[3 * e + a];
            for (int I = 0; I < ndof; I++)
                for (int J = 0; J < ndof; J++)
                    B[(size_t)v * bb + ndof * I + J] += A[(size_t)e * s * s + (size_t)(ndof * a + I) * s + ndof * a + J];
        }
    for (int v = 0; v < nnodes; v++) {   /* :79-80 */
        if (ndof == 4) inv4(B + (size_t)v * bb, inv + (size_t)v * bb);
        else if (ndof == 3) inv3(B + (size_t)v * bb, inv + (size_t)v * bb);
        else inv[v] = 1.0 / B[v];
    }
    free(B);
}

void orc_blockdiag_apply(int nnodes, int ndof, int NEQ, const int *id, const double *inv, const double *p,
                         double *z)
{
    /* BlockDiag_precond.c:12-26: p[NEQ] = 0, z[id] = inv * p[id] node by node (p == z allowed: a
     * node reads its own dofs before writing them), z[NEQ] dropped */
    double x[4];
    for (int v = 0; v < nnodes; v++) {
        const int *I = id + (size_t)v * ndof;
        const double *m = inv + (size_t)v * ndof * ndof;
        for (int j = 0; j < ndof; j++) x[j] = (I[j] < NEQ) ? p[I[j]] : 0.0;
        for (int i = 0; i < ndof; i++) {
            double t = m[ndof * i] * x[0];
            for (int j = 1; j < ndof; j++) t = t + m[ndof * i + j] * x[j];
            if (I[i] < NEQ) z[I[i]] = t;
        }
    }
}

/* ====================================================================== solvers */

static void sys_mv(const orc_system *S, const double *v, double *w)
{
    if (S->kind == 0) orc_csrmv(S->n, S->IA, S->JA, S->AA, v, w);
    else if (S->kind == 1) orc_ebemv(S->nel, S->n, S->lm, S->Perm, S->invPerm, S->A, v, w);
    else orc_ebemv_block(S->nel, S->s, S->n, S->lm, S->A, v, w);
}

static void sys_precond(const orc_system *S, double *v)
{
    if (S->precond == 1) orc_ilu0_apply(S->n, S->pIA, S->pJA, S->LU, S->Dinv, v, v);
    else if (S->precond == 2) orc_blockdiag_apply(S->bd_nnodes, S->bd_ndof, S->n, S->bd_id, S->bd_inv, v, v);
}

int orc_pgmres(const orc_system *S, int kmax, int lmax, double tol, const double *B, double *X,
               double *rho_out, int *cycles_out)
{
    const int neq = S->n;
    int i, j, k, l = 0, cont = 0;
    double **u = (double **)calloc((size_t)kmax + 1, sizeof(double *));
    double *ua = (double *)calloc((size_t)(kmax + 1) * (neq + 1), sizeof(double));
    for (i = 0; i <= kmax; i++) u[i] = ua + (size_t)i * (neq + 1);
    double *H = (double *)calloc((size_t)(kmax + 1) * (kmax + 1), sizeof(double));
#define HH(r, c) H[(size_t)(r) * (kmax + 1) + (c)]
    double *c = (double *)calloc((size_t)kmax + 1, sizeof(double));
    double *s = (double *)calloc((size_t)kmax + 1, sizeof(double));
    double *e = (double *)calloc((size_t)kmax + 1, sizeof(double));
    double *y = (double *)calloc((size_t)kmax + 1, sizeof(double));
    double *v = (double *)calloc((size_t)neq + 1, sizeof(double));
    double rho = 0;
    /* pgmres.c:39-46 */
    for (i = 0; i < neq; i++) X[i] = 0;
    const double normb = sqrt(orc_ddot(neq, B, B));
    const double eps = tol * normb;
    do {
        /* pgmres.c:54-80: restart residual (precondR is the identity for ILU/NOT) */
        i = 0;
        for (j = 0; j < kmax; j++) memset(u[j], 0, sizeof(double) * (size_t)neq);
        memcpy(v, X, sizeof(double) * (size_t)neq);
        v[neq] = 0;
        sys_mv(S, v, u[0]);
        sys_precond(S, u[0]);
        orc_daxpy(neq, -1.0, B, u[0]);
        orc_dscal(neq, -1.0, u[0]);
        e[0] = sqrt(orc_ddot(neq, u[0], u[0]));
        orc_dscal(neq, 1.0 / e[0], u[0]);
        rho = e[0];
        do {
            /* Arnoldi step (pgmres.c:87-115) */
            cont++;
            memcpy(v, u[i], sizeof(double) * (size_t)neq);
            sys_mv(S, v, u[i + 1]);
            sys_precond(S, u[i + 1]);
            for (j = 0; j <= i; j++) {
                HH(j, i) = orc_ddot(neq, u[j], u[i + 1]);
                orc_daxpy(neq, -HH(j, i), u[j], u[i + 1]);
            }
            HH(i + 1, i) = sqrt(orc_ddot(neq, u[i + 1], u[i + 1]));
            orc_dscal(neq, 1.0 / HH(i + 1, i), u[i + 1]);
            /* Givens (pgmres.c:118-156) */
            for (j = 0; j <= i - 1; j++) {
                const double h1 = c[j] * HH(j, i) + s[j] * HH(j + 1, i);
                const double h2 = -s[j] * HH(j, i) + c[j] * HH(j + 1, i);
                HH(j, i) = h1;
                HH(j + 1, i) = h2;
            }
            const double r = sqrt(HH(i, i) * HH(i, i) + HH(i + 1, i) * HH(i + 1, i));
            c[i] = HH(i, i) / r;
            s[i] = HH(i + 1, i) / r;
            HH(i, i) = r;
            HH(i + 1, i) = 0.0;
            e[i + 1] = -s[i] * e[i];
            e[i] = c[i] * e[i];
            rho = fabs(e[i + 1]);
            i++;
        } while ((rho > eps) && (i < kmax - 1));
        /* back substitution and update (pgmres.c:159-175) */
        i--;
        y[i] = e[i] / HH(i, i);
        for (j = i - 1; j >= 0; j--) {
            double soma = 0.0;
            for (k = j + 1; k <= i; k++) soma = soma + HH(j, k) * y[k];
            y[j] = (e[j] - soma) / HH(j, j);
        }
        for (j = 0; j <= i; j++)
            for (k = 0; k < neq; k++) X[k] = X[k] + u[j][k] * y[j];
        l++;
    } while ((rho > eps) && (l < lmax));
#undef HH
    free(u); free(ua); free(H); free(c); free(s); free(e); free(y); free(v);
    if (rho_out) *rho_out = rho;
    if (cycles_out) *cycles_out = l;
    return cont;
}

int orc_pcg(const orc_system *S, int itmax, double tol, const double *B, double *X)
{
    const int neq = S->n;
    int iter;
    double *P = (double *)calloc((size_t)neq + 1, sizeof(double));
    double *Q = (double *)calloc((size_t)neq + 1, sizeof(double));
    double *R = (double *)calloc((size_t)neq + 1, sizeof(double));
    double *Z = (double *)calloc((size_t)neq + 1, sizeof(double));
    /* pcg.c:31-47 */
    for (iter = 0; iter < neq; iter++) X[iter] = 0;
    memcpy(R, B, sizeof(double) * (size_t)neq);
    memcpy(Z, R, sizeof(double) * (size_t)neq);
    sys_precond(S, Z);
    memcpy(P, Z, sizeof(double) * (size_t)neq);
    double gnew = orc_ddot(neq, R, Z);
    const double r0 = gnew;
    iter = 0;
    if (!(gnew < tol * tol * r0)) {
        for (iter = 0; iter < itmax; iter++) {
            /* pcg.c:49-80 */
            sys_mv(S, P, Q);
            const double tau = orc_ddot(neq, P, Q);
            const double alpha = gnew / tau;
            orc_daxpy(neq, alpha, P, X);
            orc_daxpy(neq, -alpha, Q, R);
            memcpy(Z, R, sizeof(double) * (size_t)neq);
            sys_precond(S, Z);
            const double gold = gnew;
            gnew = orc_ddot(neq, R, Z);
            if (gnew < tol * tol * r0) break;
            const double beta = gnew / gold;
            orc_dscal(neq, beta, P);
            orc_daxpy(neq, 1.0, Z, P);
        }
    }
    free(P); free(Q); free(R); free(Z);
    return iter;
}
