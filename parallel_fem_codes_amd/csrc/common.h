// Internal declarations of libfemgpu.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/femgpu.h"

namespace fg {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define FG_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return fg::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define FG_CHECK(cond, code, ...)         \
    do {                                          \
        if (!(cond)) {                            \
            fg::set_error(__VA_ARGS__);           \
            return (code);                        \
        }                                         \
    } while (0)

#define FG_TRY(call)                  \
    do {                              \
        int r_ = (call);              \
        if (r_ != FEMGPU_OK) return r_; \
    } while (0)

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kReduceBlocks = 1024; // upper bound of blocks of the reduction kernels
constexpr int kMaxDevices = 64;     // devices a process can hold contexts on

// Typed device allocation; frees on destruction.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        if (count <= n && p) return FEMGPU_OK;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
            return FEMGPU_ERR_NOMEM;
        }
        n = count;
        return FEMGPU_OK;
    }
    int upload(const T *h, size_t count, hipStream_t s) {
        FG_TRY(alloc(count));
        if (count) FG_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return FEMGPU_OK;
    }
};

// Packed sweep plan (sweep.hip): chunks of R rows in sweep order, rows of a chunk level-sorted,
// factor entries packed in schedule order with their dependency pre-encoded (>= 0: LDS slot of a
// row of the same chunk; < 0: -(row+1) of a row of an earlier chunk, read from its mailbox).
struct PackedSweep {
    int n = 0, R = 0, nchunks = 0, ngroups = 0, depth = 0, E = 0, emax = 0;
    DevBuf<int> chunk_ptr;   // chunk -> first schedule position (nchunks+1)
    DevBuf<int> lvl_ptr;     // level -> first schedule position (depth+1); the schedule is level-sorted
    DevBuf<int> row_of;      // schedule position -> row (n)
    DevBuf<int> ent_ptr;     // schedule position -> first packed entry (n+1)
    DevBuf<int> ent_code;    // packed entry -> dependency code (E)
    DevBuf<int> ent_src;     // packed entry -> position in A's layout, to refresh values (E)
    DevBuf<int> grp_ptr;     // chunk -> first group (nchunks+1)
    DevBuf<int> grp_start;   // group -> first in-chunk position (ngroups+1)
    DevBuf<int> grp_flags;   // group -> bit 0: has a remote dependency
    DevBuf<unsigned char> exported;   // schedule position -> row read by a later chunk
    DevBuf<double> val64, dinv64;     // packed factor values / inverse pivots in schedule order
    DevBuf<float> val32, dinv32;
    DevBuf<unsigned> ctl;             // ticket, finished count, epoch
    DevBuf<unsigned long long> mbox;    // FP64: two alternating arrays of one word per row
    DevBuf<unsigned long long> mbox32;  // FP32: {epoch, value} word per row
};

// Single-workgroup streamed sweep (sweep.hip, k_sweep_stream): the level schedule cut into steps of
// at most kStreamThreads rows (a level wider than that is split), row data padded to K dependency
// entries per row in schedule order, and each dependency named by its slot in LDS: an LDS ring of
// W results (schedule position mod W; slot W holds a zero for padding, slot W+1 takes writes that
// nobody reads) followed by L long-lived slots. A dependency still in the ring when read (W >= end
// of the reading step - position of the dependency) is read from the ring; the few read later than
// that are also written, by their own row, to a long-lived slot that no other value takes until
// the step after its last reader. The plan exists when the rows have at most 16 dependencies and
// everything fits in LDS; otherwise the packed multi-workgroup sweep is used.
constexpr int kStreamThreads = 512;
constexpr int kStreamPad = 16;     // step count padded to a multiple of this (twice the largest lookahead)
constexpr int kStreamRingMax = 8192;    // ring entries at most (longer reaches take long-lived slots)
constexpr int kStreamLongMax = 4096;    // long-lived slots at most

struct StreamSweep {
    bool ok = false;
    int n = 0, nsteps = 0, steps = 0, K = 0, W = 0, L = 0, depth = 0;   // nsteps: padded, steps: real
    DevBuf<int> stepq;              // step -> first schedule position (nsteps + 2*kStreamPad + 1, tail = n)
    DevBuf<int> row_of;             // schedule position -> row (n)
    DevBuf<unsigned short> slot;    // n*K LDS slots of the dependencies
    DevBuf<unsigned short> keep;    // L > 0: n second write slots (a long-lived slot, or W + 1)
    DevBuf<int> src;                // n*K positions in A's layout, -1 for padding (value refresh)
    DevBuf<double> val64, dinv64;   // n*K factor values / n inverse pivots (upper sweep) in schedule order
    DevBuf<float> val32, dinv32;
    DevBuf<double> sink;            // kStreamThreads idle-thread results
};

// Streamed numeric factorisation (factor_stream.hip): the lower sweep's schedule and ring slots,
// with each row's values padded to kFactorRow entries and, for every (entry e, L entry t) pair, the
// index of column e in the U part of row JA[t] (4 bits; 15: absent). Results of recent rows (their
// U part and inverse pivot) live in an LDS ring of kFactorRing values per slot.
constexpr int kFactorRow = 8;    // entries per row
constexpr int kFactorU = 4;      // U entries per row kept in the ring
constexpr int kFactorRing = kFactorU + 1;
// Compact factorisation rows (at most 7 entries, 3 L, at fixed slots: L entries 0-2, the diagonal
// 3, U entries 4-6): index of the (slot e, L slot t) pair, t < min(3, e), among the row's 15 pairs.
constexpr int factor_pair_index(int e, int t) { return e == 1 ? 0 : e == 2 ? 1 + t : 3 + 3 * (e - 3) + t; }
constexpr unsigned factor_pair_bit(int e, int t) { return 1u << factor_pair_index(e, t); }
// The pair pattern of the structured triangle meshes under SYMRCM: for every row whose L entry t is
// present, the pairs of kStructAlways find column e in that row's U part and those of kStructNever
// do not (the other three pairs vary). A plan whose rows all satisfy both takes the factorisation
// variant with these pairs resolved at compile time: no select for the first, nothing for the second.
constexpr unsigned kStructAlways = factor_pair_bit(1, 0) | factor_pair_bit(2, 0) | factor_pair_bit(3, 0) |
                                   factor_pair_bit(3, 1) | factor_pair_bit(3, 2) | factor_pair_bit(5, 2);
constexpr unsigned kStructNever = factor_pair_bit(2, 1) | factor_pair_bit(5, 0) | factor_pair_bit(5, 1) |
                                  factor_pair_bit(6, 0) | factor_pair_bit(6, 1) | factor_pair_bit(6, 2);

struct StreamFactor {
    bool ok = false;
    bool compact = false;    // every row has at most 7 entries, 3 of them L and 3 U, and every L
                             // entry's row lies at most 2 047 schedule positions back
    bool structured = false; // compact, and every row fits kStructAlways / kStructNever
    DevBuf<uint4> codes;     // general, n: 32 codes of 4 bits, code (e, t) at bit 4*(4e+t)
    DevBuf<unsigned long long> rec;   // compact, n: 15 2-bit U columns of the (slot e, L slot t)
                                      // pairs, then three 11-bit distances to the L entries' rows
    DevBuf<int4> meta;       // n: {p0, len, nL, row}
    DevBuf<double> aval;     // kFactorRow*n: value e of schedule position q at e*n + q (from AA)
    DevBuf<double> out;      // factored rows in planes of n + kStreamThreads positions (idle threads
                             // write past n): general kFactorRow planes, compact 4 (pivot, U part)
    DevBuf<double> dinv;     // general, n + kStreamThreads: inverse pivots in schedule order
};

// Numeric ILU(0) plan: for every row in the lower sweep's level schedule, its CSR position, length,
// number of L entries, and its update list -- (U entry position k of a dependency row, target slot
// in this row, index t of the L entry it belongs to) -- in the serial factorisation's order.
struct FactorPlan {
    int maxrow = 0, nupd = 0, maxupd = 0;
    DevBuf<int4> rowinfo;   // per schedule position: {p0, len, nL, first update}
    DevBuf<int2> upd;       // {k, dest | t << 16}; dest == nL marks the pivot
};

}  // namespace fg

struct femgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    size_t lds_limit = 65536;      // dynamic LDS available to one workgroup
    fg::DevBuf<double> partials;   // per-block partial sums (two sets of kReduceBlocks)
    fg::DevBuf<double> res_partials;   // per-block partials of the fused residual norm
    fg::DevBuf<double> jac_partials;   // per-block partials of a product fused into the Jacobi sweeps
    double *noalloc_scratch = nullptr;   // this device's copy of the no-allocation entry point's array
    fg::DevBuf<double> scalars;    // small device scalars (dot results)
    fg::DevBuf<int> err;           // sticky device error flag (bounded waits)
    double *h_pinned = nullptr;    // pinned host mirror for scalar results
    // Solver workspaces (grown on demand, kept across calls).
    fg::DevBuf<double> gm_basis, gm_small, gm_w, gm_r;
    fg::DevBuf<int> bt_rowoff, bt_sys;        // batched solver: system row offsets, system of each row
    fg::DevBuf<double> cgs_partials;          // classical Gram-Schmidt: per-block partial dot products
    fg::DevBuf<femgpu_batch_info> wg_info;    // persistent batched solver: per-system results
    fg::DevBuf<double> bt_scal;               // batched solver: per-system norms and scales
    fg::DevBuf<float> gm_basis32, gm_w32;
    fg::DevBuf<int> gm_flags;
    fg::DevBuf<double> drv_ws;         // nonlinear/time drivers: F, X, uold
    fg::DevBuf<double> drv_ae;         // Euler driver: element matrices in element order
    volatile int *h_status = nullptr;  // pinned, device-visible solver status words
    int *d_status = nullptr;
    // captured Arnoldi steps of the FP32 inner solver (one executable graph per step), the
    // configuration they were captured for, and the private stream they are captured on
    std::vector<hipGraphExec_t> step_graphs;
    std::vector<unsigned long long> step_key;
    hipStream_t cap_stream = nullptr;
};

struct femgpu_csr {
    femgpu_ctx *ctx = nullptr;
    unsigned long long serial = 0;  // unique per matrix created in the process
    int n = 0, nnz = 0;
    std::vector<int> hIA, hJA;
    fg::DevBuf<int> IA, JA, diag;   // diag[i] = position of the diagonal entry of row i
    fg::DevBuf<double> AA;          // nnz+1 (dummy slot nnz)
    fg::DevBuf<float> AA32;         // FP32 CSR values, kept only when sell_ok && par_ok is false
    int sell_ok = 0;                    // FP32 sliced-ELL copy present (slices of 64 rows)
    int sell_wmax = 0;                  // widest slice (entries per row)
    fg::DevBuf<int> sell_ptr;           // slice -> first entry (nslices + 1); slice width = difference / 64
    fg::DevBuf<int> sell_col;           // entry k of row 64s + l at sell_ptr[s] + 64k + l (padding: the row)
    bool sell16 = false;                // the fixed-width product kernels read sell_d16 instead:
    fg::DevBuf<short> sell_d16;         // column - row as 16 bits (padding 0), when every entry fits
    fg::DevBuf<float> sell_val;         // FP32 values in that layout (padding 0), refreshed with AA32
    fg::DevBuf<double> sell_val64;      // FP64 values in that layout, kept while spmv_lanes == 0
    std::vector<int> hsell_ptr;         // host copy of sell_ptr
    // parallel FP32 ILU(0) on the sliced-ELL layout (rows of at most 16 entries)
    int par_ok = 0, par_maxw = 0;
    fg::DevBuf<int> par_uptr;           // slice -> first update (nslices + 1); updates per row = difference / 64
    fg::DevBuf<int> par_usrc;           // update u of row 64s + l at par_uptr[s] + 64u + l: SELL position of u_kj
    fg::DevBuf<unsigned char> par_ucode;   // dest entry | (L entry t) << 4; 0xff: padding
    fg::DevBuf<int> par_rinfo;          // row -> nL | len << 8
    fg::DevBuf<int> par_dpos;           // row -> SELL position of its diagonal
    fg::DevBuf<int> par_pcol;           // SELL position -> SELL position of its column's diagonal
                                        // (padding: the row's own diagonal)
    fg::DevBuf<float> par_a, par_b;     // iterates in SELL layout
    bool lu32_in_sell = false;          // the current FP32 factors are in par_a (LU32 not written)
    fg::DevBuf<double> LU, Dinv;    // ILU(0) factors in A's layout (FP32 copies live in the packed plans)
    bool factored = false;
    fg::PackedSweep plo, pup;       // packed lower / upper sweeps
    fg::StreamSweep slo, sup;       // streamed single-workgroup sweeps (used when both exist)
    int sweep_kernel = 0;           // 0: streamed when available, 1: packed
    int factor_kernel = 0;          // 0: streamed when available, 1: level-synchronous update lists
    fg::DevBuf<int> stream_l2u;     // upper-schedule position -> lower-schedule position of its row
    fg::DevBuf<int> stream_pos_up;  // row -> upper-schedule position
    fg::DevBuf<double> stream_t64;  // 2n: schedule-ordered vectors of the streamed pair
    fg::DevBuf<float> stream_t32;
    bool f32_factored = false;          // LU32/Dinv32 hold factors of the current values
    fg::DevBuf<float> LU32, Dinv32;     // FP32 factors in A's layout (mixed solver's Jacobi sweeps)
    fg::DevBuf<float> LU32b, Dinv32b;   // second iterate of the parallel factorisation
    fg::DevBuf<float> jac_a, jac_b;     // Jacobi sweep iterates
    int ell_kl = 0, ell_ku = 0;         // Jacobi sliced-ELL widest strict triangles (0: CSR kernels)
    fg::DevBuf<int> ell_sl, ell_su;     // slice pointers (64-row slices) of the lower / upper parts
    fg::DevBuf<int> ell_cl, ell_cu;     // columns, entry k of row i at sl[i/64] + 64k + i%64 (padding: column i)
    bool ell16 = false;                 // the kernels read the columns as 16-bit distances from the row:
    fg::DevBuf<unsigned short> ell_dl, ell_du;   // lower i - j, upper j - i (padding 0), when all fit
    fg::DevBuf<float> ell_vl, ell_vu;   // their FP32 factor values (padding 0), refreshed with LU32
    fg::FactorPlan fplan;           // numeric factorisation (level schedule)
    fg::FactorPlan fplan_nat;       // the same lists in row order (parallel FP32 factorisation)
    fg::StreamFactor sfac;          // streamed single-workgroup factorisation (when available)
    // FP64 parallel ILU(0) applied by Jacobi iterations (the mixed solver's preconditioner at FP64):
    // factors in the sliced-ELL layout, strict triangles in the Jacobi sliced-ELL arrays
    int prec64_jacobi = 0;              // FP64 ILU apply: 0 exact sweeps, k > 0 k Jacobi iterations
    bool par64_factored = false;        // the FP64 parallel factors hold the current values
    fg::DevBuf<double> par64_v, par64_a, par64_b, par64_d, par64_db, ell64_vl, ell64_vu, jac64_a, jac64_b;
    int spmv_lanes = 8;             // lanes per row of the FP64 CSR kernels; 0: sliced ELL (sell_val64)
    int plans = 3;                  // FEMGPU_PLANS_* built at creation
};

struct femgpu_ebe {
    femgpu_ctx *ctx = nullptr;
    int nel = 0, s = 0, ndof = 0, NEQ = 0;
    // The product is gathered per node: group g (a node with at least one unknown; groups in order
    // of the node's first unknown) sums, in element order, the ndof element rows of each element at
    // that node, so Q is written once, with no colouring, atomics or read-modify-write. Groups go
    // in slices of 64 (one lane each); slice sl owns incidence rows [iptr[sl], iptr[sl+1]) (its
    // longest list). Incidence row r, lane l: value (d, b) at ((r*ndof + d)*s + b)*64 + l, column b
    // at (r*s + b)*64 + l, so every load of a wavefront is one contiguous run.
    int ngroups = 0, nrows = 0, maxinc = 0;
    int col16 = 0;                  // 1: int16 column offsets from the group's first unknown
    bool zero_q = false;            // some unknown belongs to no element: Q is cleared first
    std::vector<int> hiptr;
    fg::DevBuf<int> iptr;           // slice -> first incidence row (nslices + 1)
    fg::DevBuf<int> len;            // group -> incidence count
    fg::DevBuf<int> eq;             // (slice*ndof + d)*64 + lane -> unknown of dof d (NEQ: prescribed)
    fg::DevBuf<int> key;            // group -> its first unknown (origin of the int16 offsets)
    fg::DevBuf<short> col16v;       // offsets (-32768: the Dirichlet sink)
    fg::DevBuf<int> col32v;         // unknowns (NEQ: the sink)
    fg::DevBuf<int> src;            // incidence slot -> e*s*s + ndof*a*s (its rows in element order), -1 padding
    fg::DevBuf<double> val;         // FP64 values in the gathered layout
    fg::DevBuf<float> val32;        // FP32 copy (mixed-precision inner solver)
    bool has_values = false;
};

struct femgpu_mesh {
    femgpu_ctx *ctx = nullptr;
    int ndof = 1;                       // 1: transport; 3: Navier-Stokes (u, v, p per node)
    int nnodes = 0, nel = 0, neq = 0, nnz = 0;
    fg::DevBuf<double> x, y, ubc;       // ubc: prescribed value per dof (ndof per node)
    fg::DevBuf<int> type, invP, ien, lm;   // type/invP per dof, lm[nel*3*ndof]
    fg::DevBuf<int> rhs_ptr, rhs_src;   // unknown -> (element*s + i) contributions, element order
    fg::DevBuf<int> aa_ptr, aa_src;     // CSR slot -> (element*s*s + ij) contributions, element order
    fg::DevBuf<int> rhs_dst, aa_dst;    // the lists are ordered by first element: list k -> unknown / slot
    fg::DevBuf<double> fe, lift;        // per element: s loads / residual entries, 9 coupling terms
    fg::DevBuf<int> mask;               // per element: Dirichlet-vertex bit mask
    fg::DevBuf<double> ke;              // element matrices when the caller passes none
    // Fused transient assembly (CSR values and residual with no element arrays in memory): rows in
    // blocks of compact mesh patches; per block its elements (sorted; two int4 each: vertices,
    // their equation numbers and the LDS slot of each vertex row in the block), per row thread
    // fz_ki incidence codes (LDS slot | CSR positions) and {row, first CSR slot, row length,
    // incidences}. Built on first use; fz_state 0 unknown, 1 built, -1 not applicable (a row longer
    // than 15 entries or with more than 16 incidences).
    int fz_state = 0, fz_nblocks = 0, fz_slots = 0, fz_maxlen = 0, fz_ki = 0;
    fg::DevBuf<int> fz_eptr;
    fg::DevBuf<int4> fz_ent;
    fg::DevBuf<unsigned> fz_inc;
    fg::DevBuf<int4> fz_row;
};

namespace fg {

// The FP64 ILU(0) preconditioner of A is ready for its current apply mode.
inline bool ilu64_ready(const femgpu_csr *A) { return A->prec64_jacobi > 0 ? A->par64_factored : A->factored; }

// Host-side integer work (structure.cpp).
int csr_init_vec(int nel, int s, int neq, int *lm, int reorder, int permute_lm, int ndofs_total, const int *type,
                 int *invP_id, std::vector<int> &IA, std::vector<int> &JA, std::vector<int> &scheme,
                 std::vector<int> &perm);
int build_packed_plan(femgpu_ctx *ctx, int n, const int *IA, const int *JA, const int *diag, bool upper,
                      PackedSweep &plan);
int packed_refresh(femgpu_ctx *ctx, femgpu_csr *A);
int build_stream_plan(femgpu_ctx *ctx, int n, const int *IA, const int *JA, const int *diag, bool upper,
                      StreamSweep &plan);
size_t stream_lds_bytes(const StreamSweep &P, size_t elem);
int build_stream_maps(femgpu_ctx *ctx, femgpu_csr *A);
int build_stream_factor(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag);
size_t factor_lds_bytes(const StreamSweep &P);
int launch_stream_factor(femgpu_ctx *ctx, femgpu_csr *A);
// natural: rows in index order (the parallel factorisation); else the lower sweep's level schedule.
int build_factor_plan(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag, bool natural,
                      FactorPlan &F);
// LDS bytes one workgroup of the packed sweep needs for T-typed values.
template <typename T>
inline size_t packed_lds_bytes(const PackedSweep &P)
{
    return sizeof(T) * (3 * static_cast<size_t>(P.R) + P.emax) + sizeof(int) * (static_cast<size_t>(P.emax) + 5 * P.R + 1);
}

// Launchers (kernels_*.hip).
int launch_spmv64(femgpu_ctx *ctx, const femgpu_csr *A, const double *x, const double *xs, double *xout,
                  double *y, const int *skip);
int launch_spmv32(femgpu_ctx *ctx, const femgpu_csr *A, const float *x, const double *xs, float *xout,
                  float *y, const int *skip);
// r = f - A x; sumsq (when non-null) receives per-block partials of ||r||^2, possible where
// residual_sumsq_blocks(A) > 0 (that many partials)
int launch_residual64(femgpu_ctx *ctx, const femgpu_csr *A, const double *x, const double *f, double *r,
                      double *sumsq = nullptr);
int residual_sumsq_blocks(const femgpu_csr *A);
int launch_sweep64(femgpu_ctx *ctx, femgpu_csr *A, const double *b, double *x, const int *skip);
int launch_sweep32(femgpu_ctx *ctx, femgpu_csr *A, const float *b, float *x, const int *skip);
int launch_ilu_factor(femgpu_ctx *ctx, femgpu_csr *A);
int launch_parilu32(femgpu_ctx *ctx, femgpu_csr *A, int sweeps);
int launch_factor_to_f32(femgpu_ctx *ctx, femgpu_csr *A);
int launch_jacobi_iteration32(femgpu_ctx *ctx, femgpu_csr *A, int upper, const float *rhs, const float *in,
                              float *out);
int blockdiag_check(const femgpu_blockdiag *B, int n);
int blockdiag_size_check(const femgpu_blockdiag *B, int NEQ, int nnodes, int ndof);
int launch_blockdiag_apply(femgpu_ctx *ctx, const femgpu_blockdiag *B, const double *p, double *z, const int *skip);
int build_jacobi_ell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag);
// dpart != nullptr: the last upper iteration also leaves jacobi_dot_blocks(A) per-block FP64 partials
// of <dv, x> (dv == nullptr: ||x||^2) in dpart; possible where jacobi_dot_blocks(A) > 0
int launch_jacobi_apply32(femgpu_ctx *ctx, femgpu_csr *A, int it, const float *b, float *x, const int *skip,
                          const float *dv = nullptr, double *dpart = nullptr);
int jacobi_dot_blocks(const femgpu_csr *A);
// The same Jacobi apply (sliced-ELL triangles only) with b = float(r / *rn) formed by its first
// lower iteration and stored in x (the application is in place on x)
int launch_jacobi_apply32_scaled(femgpu_ctx *ctx, femgpu_csr *A, int it, const double *r, const double *rn, float *x,
                                 const float *dv = nullptr, double *dpart = nullptr);
int launch_parilu64(femgpu_ctx *ctx, femgpu_csr *A, int sweeps);
int launch_jacobi_apply64(femgpu_ctx *ctx, femgpu_csr *A, int it, const double *b, double *x, const int *skip);
int build_sell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA);
int build_parilu_sell(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag);
int launch_parilu_to_csr(femgpu_ctx *ctx, femgpu_csr *A);
int launch_sell_values(femgpu_ctx *ctx, femgpu_csr *A);
bool sell64_possible(const femgpu_csr *A);   // sliced-ELL copy present and narrow enough for FP64
bool sell64_active(const femgpu_csr *A);     // the FP64 kernels read sell_val64
int set_spmv_lanes(femgpu_csr *A, int lanes);   // 0: FP64 sliced ELL (allocates and fills sell_val64)
int launch_to_f32(femgpu_ctx *ctx, const double *src, float *dst, size_t n);
int launch_ebemv(femgpu_ctx *ctx, const femgpu_ebe *E, const double *P, double *Q, const int *skip = nullptr);
int launch_ebemv32(femgpu_ctx *ctx, const femgpu_ebe *E, const float *P, float *Q, const int *skip = nullptr);
int launch_blockdiag_apply32(femgpu_ctx *ctx, const femgpu_blockdiag *B, const float *p, float *z, const int *skip);

int dot64(femgpu_ctx *ctx, int n, const double *x, const double *y, double *d_out);
// the same reduction with the per-block partial sums in `partials` (kReduceBlocks entries)
int dot64_into(femgpu_ctx *ctx, int n, const double *x, const double *y, double *partials, double *d_out);
// Element-block meshes (assembly_ns.hip): creation and the residual / CSR value gathers.
int block_mesh_create(femgpu_ctx *ctx, int ndof, int nnodes, int nel, int neq, const double *x, const double *y,
                      const int *type, const int *invP, const double *presc, const int *ien, const int *lm,
                      const int *scheme, int nnz, femgpu_mesh **out);
int launch_block_gathers(femgpu_mesh *M, const double *Ke, double *F, double *AA);
// Counting sort of (key, source) pairs by key, keeping source order inside a key (assembly maps).
// Gather map of (key, source) pairs given in source order: the keys' lists in order of their first
// source (empty keys last), dst[k] = key of list k, each list in the pairs' order.
void csr_map(int nkeys, const std::vector<std::pair<int, int>> &pairs, std::vector<int> &ptr, std::vector<int> &src,
             std::vector<int> &dst);
int check_device_error(femgpu_ctx *ctx);
int wg_threads_for(long long rows, int nsys);   // persistent batched solver: default workgroup size
int csr_autotune(femgpu_csr *A, femgpu_tune_info *info);
void release_step_graphs(femgpu_ctx *ctx);

}  // namespace fg
