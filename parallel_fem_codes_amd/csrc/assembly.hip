// Per-element SUPG assembly of the transport problems on the device.
//
// One thread per element evaluates the closed-form 3x3 SUPG(+CAU) element matrix and loads
// (Build_K_F_SUPG.c:47-168 steady, Build_M_K_R_SUPG.c:47-184 transient). Global accumulation is a
// gather instead of the reference's serial scatter: every CSR slot and every right-hand-side row
// owns the list of its element contributions in element order, so each sum is formed in exactly
// the reference's order (csr_assembly.c:24-33, F_assembly.c:5-27, Build_M_K_R_SUPG.c:173-184)
// with no atomics. Contraction is off in this file: the element arithmetic follows the
// reference's expression order operation for operation.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

struct ElemIn {
    const double *x, *y, *ubc, *u, *du;
    const int *type, *invP, *ien;
    femgpu_transport_problem pb;
    double alpha_dt;
    int neq;
};

// An element's three vertices with their coordinates, marks and state, loaded in three rounds (the
// connectivity; coordinates, marks, numbering and boundary values; the state of the free vertices),
// every load of a round issued together: loading vertex by vertex behind a branch on its mark would
// chain about ten dependent memory round trips per element. A Dirichlet vertex (type 0) takes
// ubc and du = 0; a free vertex takes u and du at its equation number. du is read when requested.
struct ElemNodes {
    int v[3], t[3];
    double X[3], Y[3], ub[3], u[3], du[3];
};

__device__ __forceinline__ void load_nodes(const ElemIn &in, int e, bool with_du, ElemNodes &N)
{
    int ip[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) N.v[a] = in.ien[3 * e + a];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int v = N.v[a];
        N.X[a] = in.x[v];
        N.Y[a] = in.y[v];
        N.t[a] = in.type[v];
        N.ub[a] = in.ubc[v];
        ip[a] = in.invP[v];
    }
    double uf[3] = {0.0, 0.0, 0.0}, duf[3] = {0.0, 0.0, 0.0};
    if (in.neq > 0) {   // else every vertex is a Dirichlet vertex
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int q = N.t[a] != 0 ? ip[a] : 0;
            uf[a] = in.u[q];
            if (with_du) duf[a] = in.du[q];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        N.u[a] = N.t[a] == 0 ? N.ub[a] : uf[a];
        N.du[a] = N.t[a] == 0 ? 0.0 : duf[a];
    }
}

struct SupgElem {
    double y23, y31, y12, x32, x13, x21, Area, invArea, tau, Be_x, Be_y, fe, feb;
    double ke[3][3];
};

__device__ __forceinline__ void supg_geometry(SupgElem &E, const double X[3], const double Y[3],
                                              const femgpu_transport_problem &pb)
{
#pragma clang fp contract(off)
    const double third = 1.0 / 3.0;
    E.y23 = Y[1] - Y[2];
    E.y31 = Y[2] - Y[0];
    E.y12 = Y[0] - Y[1];
    E.x32 = X[2] - X[1];
    E.x13 = X[0] - X[2];
    E.x21 = X[1] - X[0];
    const double TwoA = E.x21 * E.y31 - E.x13 * E.y12;
    E.Area = 0.5 * TwoA;
    E.invArea = 1.0 / E.Area;
    const double h = sqrt(fabs(TwoA));
    const double XB = third * (X[0] + X[1] + X[2]);
    const double YB = third * (Y[0] + Y[1] + Y[2]);
    E.Be_x = pb.bx[0] + pb.bx[1] * XB + pb.bx[2] * YB;
    E.Be_y = pb.by[0] + pb.by[1] * XB + pb.by[2] * YB;
    const double norm = sqrt(E.Be_x * E.Be_x + E.Be_y * E.Be_y);
    E.tau = (norm > 0) ? 0.5 * h / norm : 0;
    E.fe = pb.f0;
    E.feb = third * (E.fe + E.fe + E.fe);
}

// CAU shock-capture coefficient (CAU_ShockCapture.c:3-31; the transient form adds dueb).
__device__ __forceinline__ double cau_eu(const SupgElem &E, const femgpu_transport_problem &pb, double ue1,
                                         double ue2, double ue3, double ueb, double dueb)
{
#pragma clang fp contract(off)
    const double third = 1. / 3.;
    const double kx = pb.k, ky = pb.k, gamma = pb.gamma, invArea = E.invArea;
    const double Be_x = E.Be_x, Be_y = E.Be_y;
    const double BetaGradu = 0.5 * invArea *
                             (ue1 * (Be_x * E.y23 + Be_y * E.x32) + ue2 * (Be_x * E.y31 + Be_y * E.x13) +
                              ue3 * (Be_x * E.y12 + Be_y * E.x21));
    const double Ru = dueb + BetaGradu + gamma * ueb - E.feb;
    const double g0 = 0.5 * invArea * (E.y23 * ue1 + E.y31 * ue2 + E.y12 * ue3);
    const double g1 = 0.5 * invArea * (E.x32 * ue1 + E.x13 * ue2 + E.x21 * ue3);
    const double normGradu = sqrt(g0 * g0 + g1 * g1);
    if (!(normGradu > 0)) return 0.0;
    const double h = sqrt(fabs(2 * E.Area));
    const double Be_xp = BetaGradu * g0 / (normGradu * normGradu);
    const double Be_yp = BetaGradu * g1 / (normGradu * normGradu);
    const double ktil_p = Be_xp * Be_xp * kx + Be_yp * Be_yp * ky;
    const double normBe_p = sqrt(Be_xp * Be_xp + Be_yp * Be_yp);
    const double Pee = 0.5 * h * normBe_p * normBe_p * normBe_p / ktil_p;
    if (third * Pee < 0.7) return 0.5 * h * third * Pee * fabs(Ru) / normGradu;
    return 0.5 * h * 0.7 * fabs(Ru) / normGradu;
}

__device__ __forceinline__ void supg_ke(SupgElem &E, const femgpu_transport_problem &pb, double Eu)
{
#pragma clang fp contract(off)
    const double sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    const double kx = pb.k, ky = pb.k, gamma = pb.gamma;
    const double Be_x = E.Be_x, Be_y = E.Be_y, tau = E.tau, invArea = E.invArea, Area = E.Area;
    const double y23 = E.y23, y31 = E.y31, y12 = E.y12, x32 = E.x32, x13 = E.x13, x21 = E.x21;
    const double D12 = 0.25 * invArea * ((kx * y23 * y31) + (ky * x32 * x13));
    const double D13 = 0.25 * invArea * ((kx * y23 * y12) + (ky * x32 * x21));
    const double D23 = 0.25 * invArea * ((kx * y31 * y12) + (ky * x13 * x21));
    const double C11 = sixth * (Be_x * y23 + Be_y * x32);
    const double C12 = sixth * (Be_x * y31 + Be_y * x13);
    const double C13 = sixth * (Be_x * y12 + Be_y * x21);
    const double R11 = sixth * (gamma * Area);
    const double R12 = twelfth * (gamma * Area);
    const double Cpg12 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y31 + Be_x * Be_y * x13) + x32 * (Be_x * Be_y * y31 + Be_y * Be_y * x13));
    const double Cpg13 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) + x32 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
    const double Cpg23 = 0.25 * invArea * tau * (y31 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) + x13 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
    const double Rpg11 = sixth * tau * gamma * (Be_x * y23 + Be_y * x32);
    const double Rpg21 = sixth * tau * gamma * (Be_x * y31 + Be_y * x13);
    const double Rpg31 = sixth * tau * gamma * (Be_x * y12 + Be_y * x21);
    const double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
    const double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
    const double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);
    E.ke[0][0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
    E.ke[0][1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
    E.ke[0][2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
    E.ke[1][0] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
    E.ke[1][1] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
    E.ke[1][2] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
    E.ke[2][0] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
    E.ke[2][1] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
    E.ke[2][2] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);
}

__device__ __forceinline__ void elem_loads(const SupgElem &E, double Fe[3])
{
#pragma clang fp contract(off)
    const double sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    const double f = E.fe, fs = f + f + f;
    Fe[0] = E.Area * twelfth * (2 * f + f + f) + sixth * E.tau * fs * (E.Be_x * E.y23 + E.Be_y * E.x32);
    Fe[1] = E.Area * twelfth * (f + 2 * f + f) + sixth * E.tau * fs * (E.Be_x * E.y31 + E.Be_y * E.x13);
    Fe[2] = E.Area * twelfth * (f + f + 2 * f) + sixth * E.tau * fs * (E.Be_x * E.y12 + E.Be_y * E.x21);
}

// Steady element kernel: Ke (9), loads fe (3), lifting terms lift[2a+t] = ke[a][b_t]*ubc[v_bt] for the
// two other vertices b_t of a (ascending), mask = Dirichlet vertices.
__global__ __launch_bounds__(256) void k_elem_steady(int nel, ElemIn in, double *__restrict__ Ke,
                                                     double *__restrict__ fe, double *__restrict__ lift,
                                                     int *__restrict__ mask)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nel) return;
    ElemNodes N;
    load_nodes(in, e, false, N);
    const double *X = N.X, *Y = N.Y, *U = N.u;
    int m = 0;
    for (int a = 0; a < 3; ++a)
        if (N.t[a] == 0) m |= 1 << a;
    SupgElem E;
    supg_geometry(E, X, Y, in.pb);
    const double third = 1.0 / 3.0;
    const double ueb = third * (U[0] + U[1] + U[2]);
    const double Eu = in.pb.shock == 1 ? cau_eu(E, in.pb, U[0], U[1], U[2], ueb, 0.0) : 0.0;
    supg_ke(E, in.pb, Eu);
    double Fe[3];
    elem_loads(E, Fe);
    for (int a = 0; a < 3; ++a) {
        fe[3 * e + a] = Fe[a];
        int t = 0;
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            lift[6 * e + 2 * a + t] = E.ke[a][b] * N.ub[b];
            ++t;
        }
        for (int b = 0; b < 3; ++b) Ke[9 * static_cast<size_t>(e) + 3 * a + b] = E.ke[a][b];
    }
    mask[e] = m;
}

// Transient element evaluation (Build_M_K_R_SUPG.c:47-184): Ae = Me + alpha*dt*Ke, the loads and
// the coupling terms lift[a][b] = me[a][b]*due[b] + ke[a][b]*ue[b].
struct TransientElem {
    double Ae[3][3], lift[3][3], Fe[3];
};

__device__ __forceinline__ void transient_elem_nodes(const ElemIn &in, const ElemNodes &N, TransientElem &T)
{
#pragma clang fp contract(off)
    const double *ue = N.u, *due = N.du;
    SupgElem E;
    supg_geometry(E, N.X, N.Y, in.pb);
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    const double ueb = third * (ue[0] + ue[1] + ue[2]);
    const double dueb = third * (due[0] + due[1] + due[2]);
    const double Eu = in.pb.shock == 1 ? cau_eu(E, in.pb, ue[0], ue[1], ue[2], ueb, dueb) : 0.0;
    supg_ke(E, in.pb, Eu);
    const double C[3] = {sixth * (E.Be_x * E.y23 + E.Be_y * E.x32), sixth * (E.Be_x * E.y31 + E.Be_y * E.x13),
                         sixth * (E.Be_x * E.y12 + E.Be_y * E.x21)};
    const double sixA = sixth * E.Area, twA = twelfth * E.Area;
    elem_loads(E, T.Fe);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double me = ((a == b) ? sixA : twA) + E.tau * C[a];
            T.Ae[a][b] = me + in.alpha_dt * E.ke[a][b];
            T.lift[a][b] = me * due[b] + E.ke[a][b] * ue[b];
        }
}

__device__ __forceinline__ void transient_elem(const ElemIn &in, int e, TransientElem &T)
{
    ElemNodes N;
    load_nodes(in, e, true, N);
    transient_elem_nodes(in, N, T);
}

// Transient element kernel: Ae (9) and, per element row a, the record {fe_a, lift_a0, lift_a1,
// lift_a2} that the residual gather reads in one piece (rec[4*(3e+a) ..]). A block's 256 elements
// own one contiguous run of Ae and one of rec: each is staged in LDS and written out with 16-byte
// stores of consecutive lanes to consecutive addresses (a thread writing its own element's 72 or 96
// bytes would make every store instruction touch dozens of cache lines).
__global__ __launch_bounds__(256) void k_elem_transient(int nel, ElemIn in, double *__restrict__ Ae,
                                                        double *__restrict__ rec)
{
#pragma clang fp contract(off)
    __shared__ double2 stage[256 * 6];   // 256 records of 12 doubles, or 256 Ae of 9
    double *sd = reinterpret_cast<double *>(stage);
    const int e0 = xcd_block() * 256, e = e0 + static_cast<int>(threadIdx.x);
    if (e0 >= nel) return;   // the whole block: padding of the XCD-ordered grid
    const int nb = min(256, nel - e0);
    TransientElem T;
    if (e < nel) {
        transient_elem(in, e, T);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) sd[9 * threadIdx.x + 3 * a + b] = T.Ae[a][b];
    }
    __syncthreads();
    const int nA = 9 * nb;   // Ae + 9 e0 is 16-byte aligned (e0 is a multiple of 256)
    double *dA = Ae + 9 * static_cast<size_t>(e0);
    for (int q = threadIdx.x; q < nA / 2; q += 256) reinterpret_cast<double2 *>(dA)[q] = stage[q];
    if ((nA & 1) && threadIdx.x == 0) dA[nA - 1] = sd[nA - 1];
    __syncthreads();
    if (e < nel)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            stage[6 * threadIdx.x + 2 * a] = make_double2(T.Fe[a], T.lift[a][0]);
            stage[6 * threadIdx.x + 2 * a + 1] = make_double2(T.lift[a][1], T.lift[a][2]);
        }
    __syncthreads();
    double2 *dR = reinterpret_cast<double2 *>(rec + 12 * static_cast<size_t>(e0));
    for (int q = threadIdx.x; q < 6 * nb; q += 256) dR[q] = stage[q];
}

// Fused transient assembly: a block of kFzRows rows (a compact patch of the mesh) first evaluates
// every element that touches its rows, once, into LDS -- for each vertex row a of the element that
// belongs to the block, the slot {Ae[a][0..2], Fe_a, lift[a][0..2]} -- then its threads gather the
// rows from LDS: the CSR values of each row (each entry summed over the row's incidences in element
// order, as csr_assembly.c adds them) and the residual (Fe_a, then the three coupling terms, in
// element order, as Build_M_K_R_SUPG.c:173-182). The sums are those of the element-array path, so
// the results are bit-identical to it, and no element array is written or read back: elements on
// a patch border are evaluated once per patch that touches them instead.
// 64 rows per block: measured 125 us per assembly against 139 us with 128 rows and 193 us with 256
// on the 1001 x 1001 mesh (smaller patches re-evaluate more border elements, but a block's LDS
// shrinks with its rows and more blocks share a CU). The element evaluation is latency-bound, so
// each 64-row block runs 256 threads for it (126 us with 64, 109 us with 128, 103 us with 256).
// Element entries that carry their vertices' equation numbers (two dependent loads fewer per
// element) and incidence codes loaded ahead of the evaluation: 86 us; the gather spread over the
// block's waves (several threads per row) instead of one thread per row: 84 us. 120 VGPRs hold
// four waves per SIMD; forcing five spills registers and doubles the time (169 us). The element
// arithmetic then keeps the SIMDs busy most of the time, so the patch size is set by thread
// utilisation and border re-evaluation: a 64-row patch touches about 161 elements (1.25
// evaluations per element), a 96-row one about 233 (1.20), which 256 threads take in one pass
// (64 rows x 192 threads 78.7 us, 96 x 256 74.4 us, 128 x 320 95.3 us, 96 x 320 115.1 us).
constexpr int kFzRows = 96;        // rows per block
constexpr int kFzThreads = 256;    // threads per block: all evaluate elements, then 2 per row gather
constexpr int kFzMaxSlots = 864;   // LDS slots per block: 864 x 7 doubles = 47.3 KB at most
constexpr int kFzVals = 7;
constexpr int kFzParts = kFzThreads / kFzRows;   // gathering threads per row
static_assert(kFzParts >= 1, "a fused-assembly block needs at least one thread per row");
constexpr int kFzMaxInc = 16;      // incidences per row (elements per vertex) the fused kernel takes

// An element of a fused-assembly block, 32 bytes: ent[0] = {v0, v1, v2, q0}, ent[1] = {q1, q2,
// s0 | s1 << 16, s2 | 0xffff << 16} with the vertices v, their equation numbers q (-1: a Dirichlet
// vertex) and the LDS slot s of each vertex row that belongs to the block (0xffff: none). Storing
// the vertices' equation numbers (fixed for a mesh) lets coordinates, boundary values and state be
// loaded together right after the entry, instead of behind connectivity, marks and numbering.
__device__ __forceinline__ void load_nodes_fz(const ElemIn &in, int4 A, int4 B, ElemNodes &N)
{
    const int v[3] = {A.x, A.y, A.z}, q[3] = {A.w, B.x, B.y};
    double uf[3], duf[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        N.v[a] = v[a];
        N.t[a] = q[a] >= 0 ? 1 : 0;
        N.X[a] = in.x[v[a]];
        N.Y[a] = in.y[v[a]];
        N.ub[a] = in.ubc[v[a]];
        const int qq = q[a] >= 0 ? q[a] : 0;
        uf[a] = in.u[qq];
        duf[a] = in.du[qq];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        N.u[a] = N.t[a] == 0 ? N.ub[a] : uf[a];
        N.du[a] = N.t[a] == 0 ? 0.0 : duf[a];
    }
}

// ki: incidence rows per block in inc (the largest row's incidences, at most NI); the gathering
// threads load their row record and all their incidence codes before the element evaluation, so
// those loads overlap it instead of following the barrier one by one.
template <int MAXLEN, int NI>
__global__ __launch_bounds__(kFzThreads) void k_transient_fused(int nblocks, ElemIn in, const int *__restrict__ eptr,
                                                             const int4 *__restrict__ ent,
                                                             const unsigned *__restrict__ inc, int ki,
                                                             const int4 *__restrict__ rinfo, int nnz,
                                                             double *__restrict__ AA, double *__restrict__ R)
{
#pragma clang fp contract(off)
    extern __shared__ double slot[];
    const int w = xcd_block();
    if (w >= nblocks) return;
    const int t = static_cast<int>(threadIdx.x);
    if (w == 0 && t == 0) {
        AA[nnz] = 0.0;        // csr_assembly's dummy slot
        R[in.neq] = 0.0;      // Build_M_K_R_SUPG.c:183
    }
    // the gathering thread's row record and incidence codes (all four waves gather: thread t takes
    // row t % kFzRows), loaded before the element evaluation so that their latency overlaps it
    const int row = t % kFzRows, part = t / kFzRows;
    const int4 ri = rinfo[w * kFzRows + row];   // {row, first CSR slot, row length, incidences}
    unsigned cc[NI];
    {
        const unsigned *ip = inc + static_cast<size_t>(w) * ki * kFzRows + row;
#pragma unroll
        for (int j = 0; j < NI; ++j) cc[j] = j < ki ? ip[j * kFzRows] : 0xffffffffu;
    }
    for (int le = eptr[w] + t; le < eptr[w + 1]; le += kFzThreads) {
        const int4 A = ent[2 * le], B = ent[2 * le + 1];
        ElemNodes N;
        load_nodes_fz(in, A, B, N);
        TransientElem T;
        transient_elem_nodes(in, N, T);
        const unsigned qa[3] = {static_cast<unsigned>(B.z) & 0xffffu, static_cast<unsigned>(B.z) >> 16,
                                static_cast<unsigned>(B.w) & 0xffffu};
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (qa[a] != 0xffffu) {
                double *s = slot + kFzVals * qa[a];
                s[0] = T.Ae[a][0];
                s[1] = T.Ae[a][1];
                s[2] = T.Ae[a][2];
                s[3] = T.Fe[a];
                s[4] = T.lift[a][0];
                s[5] = T.lift[a][1];
                s[6] = T.lift[a][2];
            }
    }
    __syncthreads();
    if (ri.x < 0 || part >= kFzParts) return;   // threads past kFzParts * kFzRows only evaluate
    // Thread part p sums the row's CSR entries p, p + kFzParts, ... (each over the row's incidences
    // in element order; an element row adds to three distinct entries, so at most one of its three
    // values matches an entry, and adding +0.0 for no match leaves a sum that starts at +0.0
    // bit-identical), the last part also the residual.
    constexpr int NE = (MAXLEN + kFzParts - 1) / kFzParts;
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.0;
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        if (j < ri.w) {
            const unsigned c = cc[j];
            const double *s = slot + kFzVals * (c & 0xffffu);
            const double s0 = s[0], s1 = s[1], s2 = s[2];
            const unsigned k0 = (c >> 16) & 15u, k1 = (c >> 20) & 15u, k2 = (c >> 24) & 15u;   // 15: Dirichlet
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const unsigned kk = static_cast<unsigned>(part + kFzParts * e);
                const double v = k0 == kk ? s0 : k1 == kk ? s1 : k2 == kk ? s2 : 0.0;
                acc[e] += v;
            }
            if (part == kFzParts - 1) {
                r += s[3];
                r -= s[4];
                r -= s[5];
                r -= s[6];
            }
        }
    }
    double *out = AA + ri.y;
#pragma unroll
    for (int e = 0; e < NE; ++e)
        if (part + kFzParts * e < ri.z) out[part + kFzParts * e] = acc[e];
    if (part == kFzParts - 1) R[ri.x] = r;
}

// RHS gathers. rhs_src entries are e*3+a in element order.
__global__ void k_rhs_steady(int neq, const int *__restrict__ dst, const int *__restrict__ ptr,
                             const int *__restrict__ src, const double *__restrict__ fe,
                             const double *__restrict__ lift, const int *__restrict__ mask, double *__restrict__ F)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= neq) return;
    const int p0 = ptr[k], p1 = ptr[k + 1];
    double f = 0;
    // entry ea = 3e + a adds fe[ea], then subtracts the lifts of the element's other Dirichlet
    // vertices b != a in increasing b (lift slots 2a, 2a + 1); the first four entries' loads
    // overlap, the arithmetic is the loop's
    constexpr int U = 4;
    if (p1 > p0) {
        double fv[U], l0[U], l1[U];
        int ms[U], as[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int ea = src[p0 + u < p1 ? p0 + u : p0], e = ea / 3, a = ea % 3;
            fv[u] = fe[ea];
            ms[u] = mask[e];
            as[u] = a;
            l0[u] = lift[6 * e + 2 * a];
            l1[u] = lift[6 * e + 2 * a + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (p0 + u < p1) {
                const int a = as[u], b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;
                f += fv[u];
                if (ms[u] & (1 << b0)) f -= l0[u];
                if (ms[u] & (1 << b1)) f -= l1[u];
            }
    }
    for (int q = p0 + U; q < p1; ++q) {
        const int ea = src[q], e = ea / 3, a = ea % 3;
        f += fe[ea];
        const int m = mask[e];
        int t = 0;
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            if (m & (1 << b)) f -= lift[6 * e + 2 * a + t];
            ++t;
        }
    }
    F[dst[k]] = f;
}

// Gathers: the first kGatherUnroll entries of a list are loaded together (positions past the list's
// end clamped to its first entry, their terms not added), then summed in list order -- the same
// additions in the same order as a loop over the list, with the loads overlapped; longer lists
// finish in a loop.
constexpr int kGatherUnroll = 8;

__global__ void k_rhs_transient(int NEQ, const int *__restrict__ dst, const int *__restrict__ ptr,
                                const int *__restrict__ src, const double *__restrict__ rec, double *__restrict__ R)
{
#pragma clang fp contract(off)
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    if (k > NEQ) return;
    if (k == NEQ) {
        R[NEQ] = 0.0;
        return;
    }
    const int p0 = ptr[k], p1 = ptr[k + 1];
    double r = 0;
    if (p1 > p0) {
        double2 f0[kGatherUnroll], f1[kGatherUnroll];
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) {
            const int q = p0 + u < p1 ? p0 + u : p0;
            const double2 *t = reinterpret_cast<const double2 *>(rec + 4 * static_cast<size_t>(src[q]));
            f0[u] = t[0];   // load of row a, then its three coupling terms
            f1[u] = t[1];
        }
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u)
            if (p0 + u < p1) {
                r += f0[u].x;
                r -= f0[u].y;
                r -= f1[u].x;
                r -= f1[u].y;
            }
        for (int q = p0 + kGatherUnroll; q < p1; ++q) {
            const double2 *t = reinterpret_cast<const double2 *>(rec + 4 * static_cast<size_t>(src[q]));
            const double2 g0 = t[0], g1 = t[1];
            r += g0.x;
            r -= g0.y;
            r -= g1.x;
            r -= g1.y;
        }
    }
    R[dst[k]] = r;
}

// CSR values: AA[slot] = sum of its element entries in element order (list k for slot dst[k]);
// AA[nnz] = 0.
__global__ void k_aa_gather(int nnz, const int *__restrict__ dst, const int *__restrict__ ptr,
                            const int *__restrict__ src, const double *__restrict__ Ke, double *__restrict__ AA)
{
#pragma clang fp contract(off)
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    if (k == nnz) {
        AA[nnz] = 0.0;
        return;
    }
    AA[dst[k]] = gather_sum<kGatherUnroll>(src, Ke, ptr[k], ptr[k + 1]);
}

static ElemIn elem_in(femgpu_mesh *M, const femgpu_transport_problem *pb, const double *u, const double *du,
                      double alpha_dt)
{
    ElemIn in;
    in.x = M->x.p;
    in.y = M->y.p;
    in.ubc = M->ubc.p;
    in.u = u;
    in.du = du;
    in.type = M->type.p;
    in.invP = M->invP.p;
    in.ien = M->ien.p;
    in.pb = *pb;
    in.alpha_dt = alpha_dt;
    in.neq = M->neq;
    return in;
}

// Lists in order of first source keep a gather's threads on neighbouring elements (the element
// data they read is shared through the caches); each list's sum order is unchanged.
void csr_map(int nkeys, const std::vector<std::pair<int, int>> &pairs, std::vector<int> &ptr, std::vector<int> &src,
             std::vector<int> &dst)
{
    std::vector<int> rank(static_cast<size_t>(nkeys), -1), count(static_cast<size_t>(nkeys), 0);
    dst.clear();
    dst.reserve(static_cast<size_t>(nkeys));
    for (auto &p : pairs) {
        if (rank[p.first] < 0) {
            rank[p.first] = static_cast<int>(dst.size());
            dst.push_back(p.first);
        }
        count[p.first]++;
    }
    for (int k = 0; k < nkeys; ++k)
        if (rank[k] < 0) {
            rank[k] = static_cast<int>(dst.size());
            dst.push_back(k);
        }
    ptr.assign(static_cast<size_t>(nkeys) + 1, 0);
    for (int k = 0; k < nkeys; ++k) ptr[static_cast<size_t>(k) + 1] = ptr[k] + count[dst[k]];
    src.resize(pairs.size());
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (auto &p : pairs) src[fill[rank[p.first]]++] = p.second;
}

// The fused assembly's plan, from the mesh's device arrays (connectivity, marks, numbering and the
// CSR slot lists): rows grouped into compact patches by breadth-first growth over the mesh graph
// (up to kFzRows rows and kFzMaxSlots incidences), each patch's elements sorted, its incidences
// numbered row by row as LDS slots.
static int build_fused_plan(femgpu_mesh *M)
{
    M->fz_state = -1;
    const int nn = M->nnodes, nel = M->nel, neq = M->neq, nnz = M->nnz;
    if (neq == 0 || nel == 0 || nnz == 0) return FEMGPU_OK;
    femgpu_ctx *ctx = M->ctx;
    hipStream_t st = ctx->stream;
    std::vector<int> ien(static_cast<size_t>(nel) * 3), type(nn), invP(nn);
    std::vector<int> aptr(static_cast<size_t>(nnz) + 1), adst(nnz);
    FG_HIP(hipMemcpyAsync(ien.data(), M->ien.p, ien.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipMemcpyAsync(type.data(), M->type.p, type.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipMemcpyAsync(invP.data(), M->invP.p, invP.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipMemcpyAsync(aptr.data(), M->aa_ptr.p, aptr.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipMemcpyAsync(adst.data(), M->aa_dst.p, adst.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipStreamSynchronize(st));
    std::vector<int> asrc(aptr[nnz]);
    FG_HIP(hipMemcpyAsync(asrc.data(), M->aa_src.p, asrc.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    FG_HIP(hipStreamSynchronize(st));
    // slot of every element entry (nnz: none -- a Dirichlet row or column)
    std::vector<int> scheme(static_cast<size_t>(nel) * 9, nnz);
    for (int k = 0; k < nnz; ++k)
        for (int q = aptr[k]; q < aptr[k + 1]; ++q) scheme[asrc[q]] = adst[k];
    // incidences of every row in element order, the row's first CSR slot and length
    std::vector<int> rptr(static_cast<size_t>(neq) + 1, 0);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const int v = ien[3 * static_cast<size_t>(e) + a];
            if (type[v] == 1) ++rptr[invP[v] + 1];
        }
    for (int i = 0; i < neq; ++i) rptr[i + 1] += rptr[i];
    std::vector<int> rinc(rptr[neq]), fill(rptr.begin(), rptr.end() - 1);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const int v = ien[3 * static_cast<size_t>(e) + a];
            if (type[v] == 1) rinc[fill[invP[v]]++] = 3 * e + a;
        }
    std::vector<int> rfirst(neq, INT32_MAX), rlast(neq, -1);
    for (int i = 0; i < neq; ++i)
        for (int q = rptr[i]; q < rptr[i + 1]; ++q)
            for (int b = 0; b < 3; ++b) {
                const int sl = scheme[9 * static_cast<size_t>(rinc[q] / 3) + 3 * (rinc[q] % 3) + b];
                if (sl < nnz) {
                    rfirst[i] = std::min(rfirst[i], sl);
                    rlast[i] = std::max(rlast[i], sl);
                }
            }
    int maxlen = 0, ki = 0;
    for (int i = 0; i < neq; ++i) {
        if (rlast[i] < 0) return FEMGPU_OK;
        maxlen = std::max(maxlen, rlast[i] - rfirst[i] + 1);
        ki = std::max(ki, rptr[i + 1] - rptr[i]);
    }
    if (maxlen > 15 || ki > kFzMaxInc) return FEMGPU_OK;
    // blocks: breadth-first patches over the row graph
    std::vector<int> blk(neq, -1), order, bptr{0}, queue;
    order.reserve(neq);
    for (int seed = 0; seed < neq; ++seed) {
        if (blk[seed] >= 0) continue;
        const int b = static_cast<int>(bptr.size()) - 1;
        int rows = 0, slots = 0;
        queue.assign(1, seed);
        blk[seed] = b;
        rows = 1;
        slots = rptr[seed + 1] - rptr[seed];
        for (size_t h = 0; h < queue.size(); ++h) {
            const int i = queue[h];
            order.push_back(i);
            for (int q = rptr[i]; q < rptr[i + 1] && rows < kFzRows; ++q) {
                const int e = rinc[q] / 3;
                for (int c = 0; c < 3 && rows < kFzRows; ++c) {
                    const int v = ien[3 * static_cast<size_t>(e) + c];
                    if (type[v] != 1) continue;
                    const int j = invP[v];
                    const int sj = rptr[j + 1] - rptr[j];
                    if (blk[j] >= 0 || slots + sj > kFzMaxSlots) continue;
                    blk[j] = b;
                    queue.push_back(j);
                    ++rows;
                    slots += sj;
                }
            }
        }
        bptr.push_back(static_cast<int>(order.size()));
    }
    const int nb = static_cast<int>(bptr.size()) - 1;
    std::vector<int> eptr{0};
    std::vector<int4> ent;
    // every block's incidence codes: ki rows of kFzRows (padding 0xffffffff)
    std::vector<unsigned> inc(static_cast<size_t>(nb) * ki * kFzRows, 0xffffffffu);
    std::vector<int4> rinfo(static_cast<size_t>(nb) * kFzRows, make_int4(-1, 0, 0, 0));
    std::vector<int> qof(static_cast<size_t>(nel) * 3, -1), touched;
    int slots_max = 0;
    for (int b = 0; b < nb; ++b) {
        const int r0 = bptr[b], nr = bptr[b + 1] - r0;
        int nslot = 0;
        touched.clear();
        for (int t = 0; t < nr; ++t) {
            const int i = order[r0 + t];
            for (int q = rptr[i]; q < rptr[i + 1]; ++q) {
                qof[rinc[q]] = nslot++;
                touched.push_back(rinc[q] / 3);
            }
            rinfo[static_cast<size_t>(b) * kFzRows + t] = make_int4(i, rfirst[i], rlast[i] - rfirst[i] + 1, rptr[i + 1] - rptr[i]);
        }
        slots_max = std::max(slots_max, nslot);
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        for (int e : touched) {
            int v[3], qn[3];
            unsigned sl[3];
            for (int a = 0; a < 3; ++a) {
                v[a] = ien[3 * static_cast<size_t>(e) + a];
                qn[a] = type[v[a]] == 1 ? invP[v[a]] : -1;
                const int q = qof[3 * static_cast<size_t>(e) + a];
                sl[a] = q >= 0 ? static_cast<unsigned>(q) : 0xffffu;
            }
            ent.push_back(make_int4(v[0], v[1], v[2], qn[0]));
            ent.push_back(make_int4(qn[1], qn[2], static_cast<int>(sl[0] | (sl[1] << 16)),
                                    static_cast<int>(sl[2] | 0xffff0000u)));
        }
        eptr.push_back(static_cast<int>(ent.size() / 2));
        const size_t base = static_cast<size_t>(b) * ki * kFzRows;
        for (int t = 0; t < nr; ++t) {
            const int i = order[r0 + t];
            for (int q = rptr[i]; q < rptr[i + 1]; ++q) {
                const int ea = rinc[q], e = ea / 3, a = ea % 3;
                unsigned code = static_cast<unsigned>(qof[ea]);
                for (int c = 0; c < 3; ++c) {
                    const int sl = scheme[9 * static_cast<size_t>(e) + 3 * a + c];
                    const unsigned k = sl < nnz ? static_cast<unsigned>(sl - rfirst[i]) : 15u;
                    code |= k << (16 + 4 * c);
                }
                inc[base + static_cast<size_t>(q - rptr[i]) * kFzRows + t] = code;
            }
        }
        for (int t = 0; t < nr; ++t) {   // qof back to -1 for the next block
            const int i = order[r0 + t];
            for (int q = rptr[i]; q < rptr[i + 1]; ++q) qof[rinc[q]] = -1;
        }
    }
    if (inc.empty()) inc.push_back(0xffffffffu);
    if (ent.empty()) ent.push_back(make_int4(0, 0, 0, 0));
    FG_TRY(M->fz_eptr.upload(eptr.data(), eptr.size(), st));
    FG_TRY(M->fz_ent.upload(ent.data(), ent.size(), st));
    FG_TRY(M->fz_inc.upload(inc.data(), inc.size(), st));
    FG_TRY(M->fz_row.upload(rinfo.data(), rinfo.size(), st));
    FG_HIP(hipStreamSynchronize(st));   // the host arrays end here
    M->fz_nblocks = nb;
    M->fz_slots = slots_max;
    M->fz_maxlen = maxlen;
    M->fz_ki = ki;
    M->fz_state = 1;
    return FEMGPU_OK;
}

}  // namespace fg

using namespace fg;

// A transport mesh's device data and gather lists in M (device buffers that are large enough are
// reused).
static int mesh_setup(femgpu_mesh *M, int nnodes, int nel, int neq, const double *x, const double *y, const int *type,
                      const int *invP_id, const double *ubc, const int *ien, const int *lm, const int *scheme, int nnz)
{
    FG_CHECK(nnodes > 0 && nel >= 0 && neq >= 0 && x && y && type && invP_id && ubc && ien, FEMGPU_ERR_ARG,
             "femgpu_mesh_create: bad arguments");
    for (int v = 0; v < nnodes; ++v)
        FG_CHECK(type[v] == 0 || type[v] == 1, FEMGPU_ERR_ARG, "femgpu_mesh_create: node %d has Type %d (one rank: 0 or 1)",
                 v, type[v]);
    for (size_t q = 0; q < static_cast<size_t>(nel) * 3; ++q)
        FG_CHECK(ien[q] >= 0 && ien[q] < nnodes, FEMGPU_ERR_ARG, "femgpu_mesh_create: vertex %d out of range", ien[q]);
    femgpu_ctx *ctx = M->ctx;
    M->nnodes = nnodes;
    M->nel = nel;
    M->neq = neq;
    M->nnz = scheme ? nnz : 0;
    M->fz_state = 0;   // the fused assembly's plan is rebuilt on first use
    hipStream_t s = ctx->stream;
    int rc = FEMGPU_OK;
    auto up = [&](auto &buf, auto *h, size_t n) {
        if (rc == FEMGPU_OK) rc = buf.upload(h, n, s);
    };
    up(M->x, x, nnodes);
    up(M->y, y, nnodes);
    up(M->ubc, ubc, nnodes);
    up(M->type, type, nnodes);
    up(M->invP, invP_id, nnodes);
    up(M->ien, ien, static_cast<size_t>(nel) * 3);
    if (lm) up(M->lm, lm, static_cast<size_t>(nel) * 3);
    // unknown -> element-vertex incidences in element order (Type-1 vertices only)
    std::vector<std::pair<int, int>> pairs;
    pairs.reserve(static_cast<size_t>(nel) * 3);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const int v = ien[3 * e + a];
            if (type[v] == 1) {
                if (invP_id[v] < 0 || invP_id[v] >= neq) rc = FEMGPU_ERR_ARG;
                else pairs.emplace_back(invP_id[v], 3 * e + a);
            }
        }
    if (rc == FEMGPU_ERR_ARG) set_error("femgpu_mesh_create: invP_id out of range");
    std::vector<int> ptr, src, dst, rptr, rsrc, rdst;
    if (rc == FEMGPU_OK) {
        csr_map(neq, pairs, rptr, rsrc, rdst);
        up(M->rhs_ptr, rptr.data(), rptr.size());
        up(M->rhs_src, rsrc.data(), rsrc.size());
        up(M->rhs_dst, rdst.data(), rdst.size());
    }
    if (rc == FEMGPU_OK && scheme) {
        pairs.clear();
        for (int e = 0; e < nel; ++e)
            for (int q = 0; q < 9; ++q) {
                const int slot = scheme[9 * static_cast<size_t>(e) + q];
                if (slot < 0 || slot > nnz) rc = FEMGPU_ERR_ARG;
                else if (slot < nnz) pairs.emplace_back(slot, 9 * e + q);
            }
        if (rc == FEMGPU_ERR_ARG) set_error("femgpu_mesh_create: scheme slot out of range");
        if (rc == FEMGPU_OK) {
            csr_map(nnz, pairs, ptr, src, dst);
            up(M->aa_ptr, ptr.data(), ptr.size());
            up(M->aa_src, src.data(), src.size());
            up(M->aa_dst, dst.data(), dst.size());
        }
    }
    if (rc == FEMGPU_OK) rc = M->fe.alloc(static_cast<size_t>(nel) * 3);
    if (rc == FEMGPU_OK) rc = M->lift.alloc(static_cast<size_t>(nel) * 12);   // steady: 6, transient: 12
    if (rc == FEMGPU_OK) rc = M->mask.alloc(static_cast<size_t>(nel));
    if (rc == FEMGPU_OK) rc = M->ke.alloc(static_cast<size_t>(nel) * 9);
    if (rc == FEMGPU_OK && hipStreamSynchronize(s) != hipSuccess) rc = FEMGPU_ERR_HIP;
    return rc;
}

extern "C" int femgpu_mesh_create(femgpu_ctx *ctx, int nnodes, int nel, int neq, const double *x, const double *y,
                                  const int *type, const int *invP_id, const double *ubc, const int *ien,
                                  const int *lm, const int *scheme, int nnz, femgpu_mesh **out)
{
    FG_CHECK(ctx && out, FEMGPU_ERR_ARG, "femgpu_mesh_create: bad arguments");
    auto *M = new femgpu_mesh();
    M->ctx = ctx;
    const int rc = mesh_setup(M, nnodes, nel, neq, x, y, type, invP_id, ubc, ien, lm, scheme, nnz);
    if (rc != FEMGPU_OK) {
        delete M;
        return rc;
    }
    *out = M;
    return FEMGPU_OK;
}

extern "C" int femgpu_mesh_rebuild(femgpu_mesh *M, int nnodes, int nel, int neq, const double *x, const double *y,
                                   const int *type, const int *invP_id, const double *ubc, const int *ien,
                                   const int *lm, const int *scheme, int nnz)
{
    FG_CHECK(M && M->ndof == 1, FEMGPU_ERR_ARG, "femgpu_mesh_rebuild: needs a transport mesh");
    return mesh_setup(M, nnodes, nel, neq, x, y, type, invP_id, ubc, ien, lm, scheme, nnz);
}

extern "C" int femgpu_mesh_destroy(femgpu_mesh *M)
{
    delete M;
    return FEMGPU_OK;
}

extern "C" int femgpu_assemble_transport(femgpu_mesh *M, const femgpu_transport_problem *pb, const double *d_u,
                                         double *d_Ke, double *d_F, double *d_AA)
{
    FG_CHECK(M && pb && d_u && d_F, FEMGPU_ERR_ARG, "femgpu_assemble_transport: bad arguments");
    FG_CHECK(M->ndof == 1, FEMGPU_ERR_STATE, "femgpu_assemble_transport: not a transport mesh");
    FG_CHECK(!d_AA || M->nnz > 0 || M->aa_ptr.p, FEMGPU_ERR_STATE,
             "femgpu_assemble_transport: mesh created without a CSR slot table");
    femgpu_ctx *ctx = M->ctx;
    double *Ke = d_Ke ? d_Ke : M->ke.p;
    const ElemIn in = elem_in(M, pb, d_u, nullptr, 0.0);
    if (M->nel > 0) {
        k_elem_steady<<<(M->nel + 255) / 256, 256, 0, ctx->stream>>>(M->nel, in, Ke, M->fe.p, M->lift.p, M->mask.p);
        FG_HIP(hipGetLastError());
    }
    if (M->neq > 0) {
        k_rhs_steady<<<(M->neq + 255) / 256, 256, 0, ctx->stream>>>(M->neq, M->rhs_dst.p, M->rhs_ptr.p, M->rhs_src.p,
                                                                    M->fe.p, M->lift.p, M->mask.p, d_F);
        FG_HIP(hipGetLastError());
    }
    if (d_AA) {
        k_aa_gather<<<xcd_grid((M->nnz + 1 + 255) / 256), 256, 0, ctx->stream>>>(M->nnz, M->aa_dst.p, M->aa_ptr.p, M->aa_src.p, Ke,
                                                                        d_AA);
        FG_HIP(hipGetLastError());
    }
    return FEMGPU_OK;
}

extern "C" int femgpu_assemble_transport_transient(femgpu_mesh *M, const femgpu_transport_problem *pb, double alpha,
                                                   double dt, const double *d_u, const double *d_du, double *d_Ae,
                                                   double *d_R, double *d_AA)
{
    FG_CHECK(M && pb && d_u && d_du && d_R, FEMGPU_ERR_ARG, "femgpu_assemble_transport_transient: bad arguments");
    FG_CHECK(M->ndof == 1, FEMGPU_ERR_STATE, "femgpu_assemble_transport_transient: not a transport mesh");
    femgpu_ctx *ctx = M->ctx;
    double *Ae = d_Ae ? d_Ae : M->ke.p;
    const ElemIn in = elem_in(M, pb, d_u, d_du, alpha * dt);
    if (!d_Ae && d_AA && M->aa_ptr.p) {
        // CSR values and residual only: the fused kernel, no element arrays
        if (M->fz_state == 0) FG_TRY(build_fused_plan(M));
        if (M->fz_state == 1) {
            const size_t lds = sizeof(double) * kFzVals * static_cast<size_t>(std::max(M->fz_slots, 1));
            const int grid = xcd_grid(M->fz_nblocks);
            const int4 *ent = M->fz_ent.p, *ri = M->fz_row.p;
            const unsigned *inc = M->fz_inc.p;
            const int nb = M->fz_nblocks, ki = M->fz_ki;
            if (M->fz_maxlen <= 8 && ki <= 8)
                k_transient_fused<8, 8><<<grid, kFzThreads, lds, ctx->stream>>>(nb, in, M->fz_eptr.p, ent, inc, ki, ri,
                                                                                M->nnz, d_AA, d_R);
            else if (M->fz_maxlen <= 8)
                k_transient_fused<8, 16><<<grid, kFzThreads, lds, ctx->stream>>>(nb, in, M->fz_eptr.p, ent, inc, ki, ri,
                                                                                 M->nnz, d_AA, d_R);
            else
                k_transient_fused<16, 16><<<grid, kFzThreads, lds, ctx->stream>>>(nb, in, M->fz_eptr.p, ent, inc, ki, ri,
                                                                                  M->nnz, d_AA, d_R);
            FG_HIP(hipGetLastError());
            return FEMGPU_OK;
        }
    }
    if (M->nel > 0) {
        k_elem_transient<<<xcd_grid((M->nel + 255) / 256), 256, 0, ctx->stream>>>(M->nel, in, Ae, M->lift.p);
        FG_HIP(hipGetLastError());
    }
    k_rhs_transient<<<xcd_grid((M->neq + 1 + 255) / 256), 256, 0, ctx->stream>>>(M->neq, M->rhs_dst.p, M->rhs_ptr.p,
                                                                       M->rhs_src.p, M->lift.p, d_R);
    FG_HIP(hipGetLastError());
    if (d_AA) {
        FG_CHECK(M->aa_ptr.p, FEMGPU_ERR_STATE, "femgpu_assemble_transport_transient: no CSR slot table");
        k_aa_gather<<<xcd_grid((M->nnz + 1 + 255) / 256), 256, 0, ctx->stream>>>(M->nnz, M->aa_dst.p, M->aa_ptr.p, M->aa_src.p, Ae,
                                                                        d_AA);
        FG_HIP(hipGetLastError());
    }
    return FEMGPU_OK;
}
