// Krylov solvers on the device: pgmres (FP64, the reference's semantics), the mixed FP32-inner /
// FP64-refinement GMRES, PCG, and the fused Arnoldi kernels they share.
//
// One Arnoldi step of pgmres.c:87-156 is, on the device:
//   spmv   : w' = A (w * 1/h), storing the normalised basis vector V[i] = w * 1/h on the way
//   sweeps : w' = U^-1 L^-1 w' (level-scheduled ILU(0), in place)
//   mgs j  : w' -= h_{j-1,i} V[j-1] fused with h_{j,i} = <V[j], w'>, j = 0..i (one pass each)
//   norm   : w' -= h_{i,i} V[i] fused with ||w'||; a one-block kernel finishes the column: Givens
//            rotations, residual estimate, convergence flag (pgmres.c:118-156)
// Every launch of a step returns immediately once the cycle-done flag is set, so the host can
// enqueue one step ahead of the convergence check it reads back.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

struct GDev {
    double *H;        // (ld x ld) Hessenberg, H[r*ld + c]
    double *c, *s, *e, *y, *hnorm, *invh;   // ld each
    double *scal;     // [0] eps  [1] rho  [2] normb
    int *flags;       // [0] cycle done  [1] steps in cycle
    int ld, kmax;
    double *partials;
    int *h_status;    // host-mapped: per step {done} ; h_rho host-mapped rho per step
    double *h_rho;
    int prec_mode;    // FP32 preconditioner (host side only): 0 exact sweeps, k > 0 Jacobi iterations
    int cgs;          // FP32 inner Arnoldi with classical Gram-Schmidt (host side only)
    int graphs;       // FP32 inner Arnoldi steps replayed from captured hipGraphs (host side only)
};

__device__ __forceinline__ double &Hat(const GDev &G, int r, int c) { return G.H[r * G.ld + c]; }

// Host-mapped doubles (G.h_rho): [0, kmax) the residual estimate after each step, kMapE0 / kMapEps
// the cycle's starting residual and its target, 248..255 norms of the mixed solver.
constexpr int kMapE0 = 246, kMapEps = 247;

constexpr int kMaxBasis = 150;   // largest GMRES basis (make_gdev)

// Givens update of column i and the convergence test (pgmres.c:118-159), with h_{i+1,i} = hn.
// Called by every thread of the block: the block stages the column, the earlier rotations and
// e[i] in LDS with loads issued together, then one thread applies the rotations -- the operations
// of a serial update in the same order -- and stores the results (a serial update in global memory
// would wait for every read after a write to the same column).
__device__ void givens_column(const GDev &G, int i, double hn)
{
#pragma clang fp contract(off)
    __shared__ double col[kMaxBasis + 1], cs[kMaxBasis], sn[kMaxBasis], ei_s;
    for (int k = threadIdx.x; k <= i; k += blockDim.x) col[k] = Hat(G, k, i);
    for (int k = threadIdx.x; k < i; k += blockDim.x) {
        cs[k] = G.c[k];
        sn[k] = G.s[k];
    }
    if (threadIdx.x == 0) ei_s = G.e[i];
    __syncthreads();
    if (threadIdx.x != 0) return;
    col[i + 1] = hn;
    for (int j = 0; j <= i - 1; ++j) {
        const double h1 = cs[j] * col[j] + sn[j] * col[j + 1];
        const double h2 = -sn[j] * col[j] + cs[j] * col[j + 1];
        col[j] = h1;
        col[j + 1] = h2;
    }
    const double a = col[i], b = col[i + 1];
    const double r = sqrt(a * a + b * b);
    const double ci = a / r, si = b / r;
    G.c[i] = ci;
    G.s[i] = si;
    for (int k = 0; k < i; ++k) Hat(G, k, i) = col[k];
    Hat(G, i, i) = r;
    Hat(G, i + 1, i) = 0.0;
    const double ei = ei_s;
    G.e[i + 1] = -si * ei;
    G.e[i] = ci * ei;
    const double rho = fabs(G.e[i + 1]);
    const int steps = i + 1;
    const int done = !(rho > G.scal[0]) || (steps >= G.kmax - 1);
    G.scal[1] = rho;
    G.flags[1] = steps;
    G.flags[0] = done;
    __hip_atomic_store(&G.h_rho[i], rho, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&G.h_status[i], done ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// 16-byte vectors of the basis type (the basis and work vectors are 256-byte aligned).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int N = 4;
    static __device__ __forceinline__ float at(const float4 &v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
    static __device__ __forceinline__ void set(float4 &v, int k, float x)
    {
        if (k == 0) v.x = x; else if (k == 1) v.y = x; else if (k == 2) v.z = x; else v.w = x;
    }
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int N = 2;
    static __device__ __forceinline__ double at(const double2 &v, int k) { return k == 0 ? v.x : v.y; }
    static __device__ __forceinline__ void set(double2 &v, int k, double x)
    {
        if (k == 0) v.x = x; else v.y = x;
    }
};

// Grid reductions leave per-block partials (plain stores; the kernel boundary makes them visible)
// that a later kernel sums in a fixed order, so results do not depend on block timing and need no
// atomics or device-scope fences.

// Sum of the nb partials in a fixed order by one block of 256 threads (valid in every thread).
// Each thread adds its entries b, b + 256, ... in order, with four loads issued together.
__device__ __forceinline__ double sum_partials(const double *partials, int nb, double *red)
{
    double v = 0;
    int b = threadIdx.x;
    for (; b + 3 * 256 < nb; b += 4 * 256) {
        const double a0 = partials[b], a1 = partials[b + 256], a2 = partials[b + 512], a3 = partials[b + 768];
        v += a0;
        v += a1;
        v += a2;
        v += a3;
    }
    for (; b < nb; b += 256) v += partials[b];
    v = block_sum<256>(v, red);
    __shared__ double total;
    if (threadIdx.x == 0) total = v;
    __syncthreads();
    return total;
}

constexpr int kMgsVec = 4;

// MGS pass j of step i on w (vector type T, accumulation in FP64); passes alternate between the two
// partial sets, pass j writing set j & 1:
//   uprev != null : w -= h_{j-1,i} * uprev, where every block forms h_{j-1,i} from the nprev
//                   partials of pass j-1 in prev (the same fixed-order sum in every block) and
//                   block 0 stores it
//   unext != null : partials of h_{j,i} = <unext, w>   (unext == null: partials of ||w||^2)
template <typename T>
__global__ __launch_bounds__(256) void k_mgs(GDev G, int n, int i, int j, T *__restrict__ w,
                                             const T *__restrict__ uprev, const T *__restrict__ unext,
                                             const double *__restrict__ prev, int nprev)
{
    if (G.flags[0]) return;
    using V = Vec16<T>;
    using VT = typename V::type;
    __shared__ double red[4];
    const int nv = n / V::N, stride = gridDim.x * 256;
    // kMgsVec vector chunks per thread per step, their loads issued together; the first step's loads
    // are issued before the previous pass's partial sums are combined into h, so the two overlap
    VT wv[kMgsVec], pv[kMgsVec], uv[kMgsVec];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < kMgsVec; ++u) {
            const int k = k0 + u * stride;
            if (k < nv) {
                wv[u] = reinterpret_cast<const VT *>(w)[k];
                if (uprev) pv[u] = reinterpret_cast<const VT *>(uprev)[k];
                if (unext) uv[u] = reinterpret_cast<const VT *>(unext)[k];
            }
        }
    };
    const int kfirst = blockIdx.x * 256 + threadIdx.x;
    load(kfirst);
    T h = T(0);
    if (uprev) {
        const double hp = sum_partials(prev, nprev, red);
        if (blockIdx.x == 0 && threadIdx.x == 0) Hat(G, j - 1, i) = hp;
        h = static_cast<T>(hp);
    }
    double acc = 0;
    for (int k0 = kfirst; k0 < nv; k0 += kMgsVec * stride) {
        if (k0 != kfirst) load(k0);
#pragma unroll
        for (int u = 0; u < kMgsVec; ++u) {
            const int k = k0 + u * stride;
            if (k >= nv) continue;
            if (uprev) {
#pragma unroll
                for (int c = 0; c < V::N; ++c) V::set(wv[u], c, V::at(wv[u], c) - h * V::at(pv[u], c));
                reinterpret_cast<VT *>(w)[k] = wv[u];
            }
            const VT &uu = unext ? uv[u] : wv[u];
#pragma unroll
            for (int c = 0; c < V::N; ++c)
                acc += static_cast<double>(V::at(uu, c)) * static_cast<double>(V::at(wv[u], c));
        }
    }
    if (blockIdx.x == 0)
        for (int k = nv * V::N + threadIdx.x; k < n; k += 256) {
            T wv = w[k];
            if (uprev) {
                wv = wv - h * uprev[k];
                w[k] = wv;
            }
            const T u = unext ? unext[k] : wv;
            acc += static_cast<double>(u) * static_cast<double>(wv);
        }
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) G.partials[(j & 1) * kReduceBlocks + blockIdx.x] = acc;
}

// After the norm pass (j = i + 1): h_{i+1,i} = ||w||, then the Givens update of column i.
__global__ __launch_bounds__(256) void k_mgs_finish(GDev G, int nb, int i)
{
    if (G.flags[0]) return;
    __shared__ double red[4];
    const double v = sum_partials(G.partials + ((i + 1) & 1) * kReduceBlocks, nb, red);
    const double hn = sqrt(v);
    if (threadIdx.x == 0) {
        G.hnorm[i + 1] = hn;
        G.invh[i + 1] = 1.0 / hn;
    }
    givens_column(G, i, hn);
}

// Classical Gram-Schmidt for the FP32 inner solver: all coefficients h_{j,i} = <V[j], w> from one
// pass over w and the basis (per-block partial sums, FP64), one kernel that reduces them (a block
// per coefficient), then one pass w -= sum_j h_{j,i} V[j] (j ascending) fused with the partials of
// ||w||^2. The coefficients come from the same w (not the successively updated one of MGS), which
// the inner solve of the refinement tolerates; the outer FP64 loop verifies the answer.
constexpr int kCgsBlocks = kReduceBlocks, kCgsChunk = 8;

__global__ __launch_bounds__(256) void k_cgs_dots(GDev G, int n, int i, const float *__restrict__ w,
                                                  const float *__restrict__ V, size_t ldv, double *__restrict__ part)
{
    if (G.flags[0]) return;
    __shared__ double red[4 * kCgsChunk];
    const int nv = n / 4;
    for (int j0 = 0; j0 <= i; j0 += kCgsChunk) {
        double acc[kCgsChunk];
#pragma unroll
        for (int c = 0; c < kCgsChunk; ++c) acc[c] = 0.0;
        for (int k = blockIdx.x * 256 + threadIdx.x; k < nv; k += gridDim.x * 256) {
            const float4 wv = reinterpret_cast<const float4 *>(w)[k];
            float4 v[kCgsChunk];
#pragma unroll
            for (int c = 0; c < kCgsChunk; ++c)
                if (j0 + c <= i) v[c] = reinterpret_cast<const float4 *>(V + (j0 + c) * ldv)[k];
#pragma unroll
            for (int c = 0; c < kCgsChunk; ++c)
                if (j0 + c <= i)
                    acc[c] += static_cast<double>(v[c].x) * wv.x + static_cast<double>(v[c].y) * wv.y +
                              static_cast<double>(v[c].z) * wv.z + static_cast<double>(v[c].w) * wv.w;
        }
        if (blockIdx.x == 0)
            for (int k = nv * 4 + threadIdx.x; k < n; k += 256)
#pragma unroll
                for (int c = 0; c < kCgsChunk; ++c)
                    if (j0 + c <= i) acc[c] += static_cast<double>(V[(j0 + c) * ldv + k]) * w[k];
        // all coefficients of the chunk reduced together: wave sums, one barrier, then thread c
        // adds the four wave sums of coefficient c in wave order
#pragma unroll
        for (int c = 0; c < kCgsChunk; ++c) acc[c] = wave_sum(acc[c]);
        __syncthreads();
        if (threadIdx.x % 64 == 0)
#pragma unroll
            for (int c = 0; c < kCgsChunk; ++c) red[c * 4 + threadIdx.x / 64] = acc[c];
        __syncthreads();
        if (threadIdx.x < kCgsChunk && j0 + threadIdx.x <= i)
            part[(j0 + threadIdx.x) * kCgsBlocks + blockIdx.x] =
                ((red[threadIdx.x * 4] + red[threadIdx.x * 4 + 1]) + red[threadIdx.x * 4 + 2]) + red[threadIdx.x * 4 + 3];
    }
}

// Block j: h_{j,i} = the fixed-order sum of the nb partials of coefficient j.
__global__ __launch_bounds__(256) void k_cgs_reduce(GDev G, int nb, int i, const double *__restrict__ part)
{
    if (G.flags[0]) return;
    __shared__ double red[4];
    const int j = blockIdx.x;
    const double v = sum_partials(part + j * kCgsBlocks, nb, red);
    if (threadIdx.x == 0) Hat(G, j, i) = v;
}

// w -= sum_j h_{j,i} V[j] (FP32, j ascending; loads of a chunk of basis vectors issued together) and
// the partials of ||w||^2.
__global__ __launch_bounds__(256) void k_cgs_update(GDev G, int n, int i, float *__restrict__ w,
                                                    const float *__restrict__ V, size_t ldv, double *__restrict__ part)
{
    if (G.flags[0]) return;
    __shared__ double red[4];
    __shared__ float h[160];
    for (int j = threadIdx.x; j <= i; j += 256) h[j] = static_cast<float>(Hat(G, j, i));
    __syncthreads();
    double acc = 0.0;
    const int nv = n / 4;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < nv; k += gridDim.x * 256) {
        float4 wv = reinterpret_cast<const float4 *>(w)[k];
        for (int j0 = 0; j0 <= i; j0 += kCgsChunk) {
            float4 v[kCgsChunk];
#pragma unroll
            for (int c = 0; c < kCgsChunk; ++c)
                if (j0 + c <= i) v[c] = reinterpret_cast<const float4 *>(V + (j0 + c) * ldv)[k];
#pragma unroll
            for (int c = 0; c < kCgsChunk; ++c)
                if (j0 + c <= i) {
                    const float hj = h[j0 + c];
                    wv.x -= hj * v[c].x;
                    wv.y -= hj * v[c].y;
                    wv.z -= hj * v[c].z;
                    wv.w -= hj * v[c].w;
                }
        }
        reinterpret_cast<float4 *>(w)[k] = wv;
        acc += static_cast<double>(wv.x) * wv.x + static_cast<double>(wv.y) * wv.y +
               static_cast<double>(wv.z) * wv.z + static_cast<double>(wv.w) * wv.w;
    }
    if (blockIdx.x == 0)
        for (int k = nv * 4 + threadIdx.x; k < n; k += 256) {
            float wv = w[k];
            for (int j = 0; j <= i; ++j) wv -= h[j] * V[j * ldv + k];
            w[k] = wv;
            acc += static_cast<double>(wv) * wv;
        }
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) part[(i + 1) * kCgsBlocks + blockIdx.x] = acc;
}

// h_{i+1,i} = ||w||, then the Givens update of column i.
__global__ __launch_bounds__(256) void k_cgs_finish(GDev G, int nb, int i, const double *__restrict__ part)
{
    if (G.flags[0]) return;
    __shared__ double red[4];
    const double v = sum_partials(part + (i + 1) * kCgsBlocks, nb, red);
    const double hn = sqrt(v);
    if (threadIdx.x == 0) {
        G.hnorm[i + 1] = hn;
        G.invh[i + 1] = 1.0 / hn;
    }
    givens_column(G, i, hn);
}

// Start of a cycle from the residual vector w (already M^-1 (B - A X) or B itself):
//   negate = 1: w = B - w (pgmres.c:73-74), then e0 = ||w||, invh[0] = 1/e0, rho = e0, flags reset
template <typename T>
__global__ __launch_bounds__(256) void k_cycle_start(GDev G, int n, T *__restrict__ w, const T *__restrict__ B,
                                                     int negate)
{
    __shared__ double red[4];
    double acc = 0;
    const int stride = gridDim.x * 256;
    for (int k0 = blockIdx.x * 256 + threadIdx.x; k0 < n; k0 += 4 * stride) {   // four loads in flight
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * stride;
            v[u] = k < n ? w[k] : T(0);
            if (negate && k < n) v[u] = B[k] - v[u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * stride;
            if (k >= n) continue;
            if (negate) w[k] = v[u];
            acc += static_cast<double>(v[u]) * static_cast<double>(v[u]);
        }
    }
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) G.partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_cycle_start_finish(GDev G, int nb, int first, double tol,
                                                            const double *__restrict__ part)
{
    __shared__ double red[4];
    for (int q = threadIdx.x; q < G.ld * G.ld; q += 256) G.H[q] = 0.0;   // the cycle's Hessenberg starts at zero
    const double v = sum_partials(part, nb, red);
    if (threadIdx.x == 0) {
        const double e0 = sqrt(v);
        if (first) {   // pgmres.c:42-46: ||B|| and eps
            G.scal[2] = e0;
            G.scal[0] = tol * e0;
        }
        G.e[0] = e0;
        G.hnorm[0] = e0;
        G.invh[0] = 1.0 / e0;
        G.scal[1] = e0;
        G.flags[0] = 0;
        G.flags[1] = 0;
        // the cycle's starting residual and target for the host's look-ahead
        __hip_atomic_store(&G.h_rho[kMapE0], e0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&G.h_rho[kMapEps], G.scal[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Back substitution H y = e over the first i + 1 columns (pgmres.c:159-170) into ys (LDS), the
// same operations in the same order as the serial loop; called by every thread of the block. For
// bases up to kStageBasis the block first stages the triangle and e in LDS with loads issued
// together, so the serial loop of thread 0 does not wait on global memory per step.
constexpr int kStageBasis = 32;

__device__ void back_substitute(const GDev &G, int i, double *ys)
{
#pragma clang fp contract(off)
    __shared__ double hs[kStageBasis * kStageBasis], es[kStageBasis];
    if (i < kStageBasis) {
        const int m = i + 1;
        for (int q = threadIdx.x; q < m * m; q += blockDim.x) {
            const int r = q / m, c = q - r * m;
            if (c >= r) hs[q] = Hat(G, r, c);
        }
        for (int q = threadIdx.x; q < m; q += blockDim.x) es[q] = G.e[q];
        __syncthreads();
        if (threadIdx.x == 0) {
            ys[i] = es[i] / hs[i * m + i];
            for (int j = i - 1; j >= 0; --j) {
                double soma = 0.0;
                for (int k = j + 1; k <= i; ++k) soma = soma + hs[j * m + k] * ys[k];
                ys[j] = (es[j] - soma) / hs[j * m + j];
            }
        }
    } else if (threadIdx.x == 0) {
        ys[i] = G.e[i] / Hat(G, i, i);
        for (int j = i - 1; j >= 0; --j) {
            double soma = 0.0;
            for (int k = j + 1; k <= i; ++k) soma = soma + Hat(G, j, k) * ys[k];
            ys[j] = (G.e[j] - soma) / Hat(G, j, j);
        }
    }
}

// Back substitution (pgmres.c:159-170) by every block into LDS, then X += scale * sum_j V[j] y[j]
// in the reference's per-entry order (pgmres.c:173-175).
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_cycle_finish(GDev G, int n, const T *__restrict__ V, size_t ldv,
                                                      TX *__restrict__ X, double scale, int assign)
{
#pragma clang fp contract(off)
    __shared__ double ys[kMaxBasis + 1];
    const int i = G.flags[1] - 1;
    if (i < 0) return;
    back_substitute(G, i, ys);
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int j = 0; j <= i; ++j) G.y[j] = ys[j];
    __syncthreads();
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        // assign: X was not initialised (x0 = 0): the sums start from 0 instead of X
        if (scale == 1.0) {
            TX xv = assign ? TX(0) : X[k];
            for (int j = 0; j <= i; ++j) xv = xv + static_cast<TX>(V[j * ldv + k]) * static_cast<TX>(ys[j]);
            X[k] = xv;
        } else {
            double d = 0;
            for (int j = 0; j <= i; ++j) d = d + static_cast<double>(V[j * ldv + k]) * ys[j];
            X[k] = static_cast<TX>((assign ? 0.0 : static_cast<double>(X[k])) + scale * d);
        }
    }
}

// Scaled FP32 copy of the FP64 residual: r32 = r * (1/rn).
__global__ void k_scale_to_f32(int n, const double *__restrict__ r, const double *__restrict__ rn,
                               float *__restrict__ r32)
{
    const double inv = 1.0 / *rn;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        r32[k] = static_cast<float>(r[k] * inv);
}

// r = F - r (the FP64 residual of the EBE operator: r holds A x on entry)
__global__ void k_rsub(int n, const double *__restrict__ F, double *__restrict__ r)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) r[k] = F[k] - r[k];
}

__global__ void k_sqrt_inplace(double *v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *v = sqrt(*v);
}

// ||x||: per-block partial sums of squares, then one block sums them in a fixed order and stores
// the square root both in device memory (for later kernels) and in a host-mapped word (read by the
// host after the stream synchronises, with no copy).
__global__ __launch_bounds__(256) void k_sumsq(int n, const double *__restrict__ x, double *__restrict__ partials)
{
    __shared__ double red[4];
    double acc = 0.0;
    const int stride = gridDim.x * 256;
    for (int k0 = blockIdx.x * 256 + threadIdx.x; k0 < n; k0 += 4 * stride) {   // four loads in flight
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = k0 + u * stride < n ? x[k0 + u * stride] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u] * v[u];
    }
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_norm_finish(int nb, const double *__restrict__ partials, double *d_out,
                                                     double *h_out)
{
    __shared__ double red[4];
    const double v = sum_partials(partials, nb, red);
    if (threadIdx.x == 0) {
        const double r = sqrt(v);
        *d_out = r;
        __hip_atomic_store(h_out, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---- batched systems: K independent systems stored as one block-diagonal matrix, system s on rows
// [off[s], off[s+1]).

// System of every row (binary search over the K+1 offsets).
__global__ void k_sys_of_row(int n, int K, const int *__restrict__ off, int *__restrict__ sys)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int lo = 0, hi = K;   // off[lo] <= i < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        sys[i] = lo;
    }
}

// ||v_s||^2 of every system, one block per system (fixed order: deterministic).
__global__ __launch_bounds__(256) void k_seg_norm2(const int *__restrict__ off, const double *__restrict__ v,
                                                   double *__restrict__ out)
{
    __shared__ double red[4];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += 256) acc += v[i] * v[i];
    acc = block_sum<256>(acc, red);
    if (threadIdx.x == 0) out[s] = acc;
}

// r32 = r * scale[system]: every active system's residual normalised to unit norm, finished
// systems zeroed (their blocks of every Krylov vector then stay exactly zero).
__global__ void k_scale_to_f32_seg(int n, const int *__restrict__ sys, const double *__restrict__ r,
                                   const double *__restrict__ scale, float *__restrict__ r32)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        r32[k] = static_cast<float>(r[k] * scale[sys[k]]);
}

// X += rn[system] * (V y): the cycle's correction of each system scaled back by its residual norm.
__global__ __launch_bounds__(256) void k_cycle_finish_seg(GDev G, int n, const float *__restrict__ V, size_t ldv,
                                                          double *__restrict__ X, const int *__restrict__ sys,
                                                          const double *__restrict__ rn)
{
#pragma clang fp contract(off)
    __shared__ double ys[kMaxBasis + 1];
    const int i = G.flags[1] - 1;
    if (i < 0) return;
    back_substitute(G, i, ys);
    __syncthreads();
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        double d = 0;
        for (int j = 0; j <= i; ++j) d = d + static_cast<double>(V[j * ldv + k]) * ys[j];
        X[k] = X[k] + rn[sys[k]] * d;
    }
}

static int grid_for(int n) { return std::max(1, std::min(kReduceBlocks, (n + 1023) / 1024)); }
// reductions whose threads take four elements per step
static int grid4_for(int n) { return std::max(1, std::min(kReduceBlocks, (n + 4095) / 4096)); }

// ||x|| into d_out and the host-mapped slot of the context (slots 248..255 of the mapped doubles,
// past the per-step residual words), with no wait; the slot is valid once the stream has passed
// the launches.
static int norm64_async(femgpu_ctx *ctx, int n, const double *x, double *d_out, int slot, volatile double *&hmap)
{
    const int grid = grid4_for(n);
    double *dmap = reinterpret_cast<double *>(ctx->d_status + 256) + 248 + slot;
    hmap = reinterpret_cast<volatile double *>(ctx->h_status + 256) + 248 + slot;
    k_sumsq<<<grid, 256, 0, ctx->stream>>>(n, x, ctx->partials.p);
    k_norm_finish<<<1, 256, 0, ctx->stream>>>(grid, ctx->partials.p, d_out, dmap);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

// ||r|| from nb per-block partials of ||r||^2 already in `partials`, into d_out and the mapped slot;
// waits for the stream and returns it.
static int norm64_partials_sync(femgpu_ctx *ctx, int nb, const double *partials, double *d_out, int slot,
                                double &host)
{
    double *dmap = reinterpret_cast<double *>(ctx->d_status + 256) + 248 + slot;
    volatile double *hmap = reinterpret_cast<volatile double *>(ctx->h_status + 256) + 248 + slot;
    k_norm_finish<<<1, 256, 0, ctx->stream>>>(nb, partials, d_out, dmap);
    FG_HIP(hipGetLastError());
    FG_HIP(hipStreamSynchronize(ctx->stream));
    host = *hmap;
    return FEMGPU_OK;
}

// The same, then waits for the stream and returns the norm.
static int norm64_sync(femgpu_ctx *ctx, int n, const double *x, double *d_out, int slot, double &host)
{
    volatile double *hmap = nullptr;
    FG_TRY(norm64_async(ctx, n, x, d_out, slot, hmap));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    host = *hmap;
    return FEMGPU_OK;
}

// The MGS passes of Arnoldi step i (j = 0..i, then the norm pass) and the column's finish. nfused > 0:
// the kernel that produced w already left nfused partials of <V[0], w> in fpart (pass 0 is skipped).
template <typename T>
static void mgs_step(femgpu_ctx *ctx, const GDev &G, int n, int i, T *wn, const T *V, size_t ldv, int nfused = 0,
                     const double *fpart = nullptr)
{
    const int nv = n / Vec16<T>::N;
    const int grid = std::max(1, std::min(kReduceBlocks, (nv + 256 * kMgsVec - 1) / (256 * kMgsVec)));
    for (int j = nfused > 0 ? 1 : 0; j <= i + 1; ++j) {
        const T *uprev = j ? V + (j - 1) * ldv : nullptr;
        const T *unext = j <= i ? V + j * ldv : nullptr;
        const bool from_fused = j == 1 && nfused > 0;
        const double *prev = from_fused ? fpart : G.partials + ((j - 1) & 1) * kReduceBlocks;
        k_mgs<T><<<grid, 256, 0, ctx->stream>>>(G, n, i, j, wn, uprev, unext, prev, from_fused ? nfused : grid);
    }
    k_mgs_finish<<<1, 256, 0, ctx->stream>>>(G, grid, i);
}

// Start of a cycle from w: ||w|| (nfused > 0: from the nfused partials of ||w||^2 that the kernel
// producing w left in fpart), the Hessenberg reset and the first column's scalars.
template <typename T>
static void cycle_start(femgpu_ctx *ctx, const GDev &G, int n, T *w, const T *B, int negate, int first, double tol,
                        int nfused = 0, const double *fpart = nullptr)
{
    if (nfused > 0 && !negate) {
        k_cycle_start_finish<<<1, 256, 0, ctx->stream>>>(G, nfused, first, tol, fpart);
        return;
    }
    const int grid = grid4_for(n);
    k_cycle_start<T><<<grid, 256, 0, ctx->stream>>>(G, n, w, B, negate);
    k_cycle_start_finish<<<1, 256, 0, ctx->stream>>>(G, grid, first, tol, G.partials);
}

// ------------------------------------------------------------------------------- workspace
template <typename T>
static int ensure_ws(femgpu_ctx *ctx, int n, int kmax, DevBuf<T> &basis, DevBuf<T> &wbuf, size_t &ldv)
{
    ldv = (static_cast<size_t>(n) + 63) / 64 * 64;
    FG_TRY(basis.alloc(ldv * (kmax + 1)));
    FG_TRY(wbuf.alloc(ldv * 3));
    return FEMGPU_OK;
}

static int make_gdev(femgpu_ctx *ctx, int kmax, GDev &G)
{
    const int ld = kmax + 1;
    FG_CHECK(kmax >= 2 && kmax <= kMaxBasis, FEMGPU_ERR_ARG, "GMRES basis size %d outside [2, %d]", kmax, kMaxBasis);
    FG_TRY(ctx->gm_small.alloc(static_cast<size_t>(ld) * ld + 6 * ld + 8));
    FG_TRY(ctx->gm_flags.alloc(8));
    double *base = ctx->gm_small.p;
    G.H = base;
    G.c = base + ld * ld;
    G.s = G.c + ld;
    G.e = G.s + ld;
    G.y = G.e + ld;
    G.hnorm = G.y + ld;
    G.invh = G.hnorm + ld;
    G.scal = G.invh + ld;
    G.flags = ctx->gm_flags.p;
    G.ld = ld;
    G.kmax = kmax;
    G.partials = ctx->partials.p;
    G.h_status = ctx->d_status;
    G.h_rho = reinterpret_cast<double *>(ctx->d_status + 256);
    return FEMGPU_OK;   // H is cleared by every cycle start (k_cycle_start_finish)
}

// Operator + preconditioner application w_out = M^-1 A (x * xs), storing x*xs into xout.
static int apply_op64(femgpu_ctx *ctx, const femgpu_system *S, const double *x, const double *xs, double *xout,
                      double *w_out, const int *skip);

// One Arnoldi cycle (steps until the device flag says done), FP64 or FP32 vectors. Returns the
// steps the device took in the cycle.
// FP32 preconditioner of the inner solver: the exact level-scheduled sweeps on the FP32 copy of the
// ILU(0) factors, or mode Jacobi iterations per triangle on the FP32 factors in A's layout.
// dpart != nullptr (Jacobi mode only): the last upper iteration leaves the partials of <dv, x>
// (dv == nullptr: ||x||^2) in dpart.
static int prec32(femgpu_ctx *ctx, femgpu_csr *A, int mode, const float *b, float *x, const int *skip,
                  const float *dv = nullptr, double *dpart = nullptr)
{
    if (mode == 0) {
        FG_CHECK(!dpart, FEMGPU_ERR_ARG, "prec32: no fused product with the exact sweeps");
        return launch_sweep32(ctx, A, b, x, skip);
    }
    return launch_jacobi_apply32(ctx, A, mode, b, x, skip, dv, dpart);
}

// Partials of a product the inner FP32 preconditioner can fuse (0: none): the Jacobi mode on the
// sliced-ELL triangles.
static int prec32_dot_blocks(const femgpu_system *S, int mode)
{
    return S->csr && S->ilu && mode > 0 ? jacobi_dot_blocks(S->ilu) : 0;
}

template <typename T>
static int run_cycle(femgpu_ctx *ctx, const femgpu_system *S, GDev &G, int n, T *V, size_t ldv, T *wA, T *wB,
                     int kmax, int *steps_out);

template <>
int run_cycle<double>(femgpu_ctx *ctx, const femgpu_system *S, GDev &G, int n, double *V, size_t ldv, double *wA,
                      double *wB, int kmax, int *steps_out)
{
    hipEvent_t ev[2];
    FG_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    FG_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    double *wc = wA, *wn = wB;
    int rc = FEMGPU_OK, steps = 0;
    for (int i = 0; i < kmax - 1; ++i) {
        ctx->h_status[i] = 0;
        rc = apply_op64(ctx, S, wc, G.invh + i, V + i * ldv, wn, G.flags);
        if (rc) break;
        mgs_step<double>(ctx, G, n, i, wn, V, ldv);
        if (hipGetLastError() != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
        if (hipEventRecord(ev[i & 1], ctx->stream) != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
        std::swap(wc, wn);
        steps = i + 1;
        if (i >= 1) {
            if (hipEventSynchronize(ev[(i - 1) & 1]) != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
            if (ctx->h_status[i - 1]) break;   // step i was enqueued speculatively and drains
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    *steps_out = steps;
    if (rc) set_error("pgmres: launch failed");
    return rc;
}

template <typename T>
__global__ void k_scale_copy(int n, const T *__restrict__ x, const double *__restrict__ xs, T *__restrict__ out,
                             const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const T s = static_cast<T>(*xs);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = x[k] * s;
}

// FP32 operator and preconditioner of the inner solver: w = M^-1 A (x * xs), with x * xs stored
// into xout (the basis vector). CSR: the sliced-ELL SpMV with the scaling fused, then the ILU(0)
// sweeps; EBE: the scaled copy, the gathered element product, then the nodal block-diagonal blocks.
static int op_prec32(femgpu_ctx *ctx, const femgpu_system *S, int prec_mode, const float *x, const double *xs,
                     float *xout, float *w, const int *skip, const float *dv = nullptr, double *dpart = nullptr)
{
    if (S->csr) {
        FG_TRY(launch_spmv32(ctx, S->csr, x, xs, xout, w, skip));
        return prec32(ctx, S->ilu, prec_mode, w, w, skip, dv, dpart);
    }
    k_scale_copy<float><<<grid_for(S->ebe->NEQ) * 4, 256, 0, ctx->stream>>>(S->ebe->NEQ, x, xs, xout, skip);
    FG_HIP(hipGetLastError());
    FG_TRY(launch_ebemv32(ctx, S->ebe, xout, w, skip));
    return launch_blockdiag_apply32(ctx, S->bdiag, w, w, skip);
}

// The launches of Arnoldi step i of the FP32 inner solver: w_next = M^-1 A (v_i), v_i = w_cur / h
// stored on the way, then the orthogonalization and the Givens update of column i. With the Jacobi
// preconditioner and MGS, the last Jacobi iteration also forms the partials of <V[0], w_next>
// (the first MGS pass), in ctx->jac_partials.
static int step_kernels32(femgpu_ctx *ctx, const femgpu_system *S, const GDev &G, int n, int i, float *wc,
                          float *wn, float *V, size_t ldv)
{
    const int nf = G.cgs ? 0 : prec32_dot_blocks(S, G.prec_mode);
    FG_TRY(op_prec32(ctx, S, G.prec_mode, wc, G.invh + i, V + i * ldv, wn, G.flags, nf ? V : nullptr,
                     nf ? ctx->jac_partials.p : nullptr));
    if (G.cgs) {
        const int nb = std::max(1, std::min(kCgsBlocks, (n / 4 + 255) / 256));
        k_cgs_dots<<<nb, 256, 0, ctx->stream>>>(G, n, i, wn, V, ldv, ctx->cgs_partials.p);
        k_cgs_reduce<<<i + 1, 256, 0, ctx->stream>>>(G, nb, i, ctx->cgs_partials.p);
        k_cgs_update<<<nb, 256, 0, ctx->stream>>>(G, n, i, wn, V, ldv, ctx->cgs_partials.p);
        k_cgs_finish<<<1, 256, 0, ctx->stream>>>(G, nb, i, ctx->cgs_partials.p);
    } else {
        mgs_step<float>(ctx, G, n, i, wn, V, ldv, nf, ctx->jac_partials.p);
    }
    return hipGetLastError() == hipSuccess ? FEMGPU_OK : FEMGPU_ERR_HIP;
}

}  // namespace fg

void fg::release_step_graphs(femgpu_ctx *ctx)
{
    for (hipGraphExec_t &x : ctx->step_graphs)
        if (x) (void)hipGraphExecDestroy(x);
    ctx->step_graphs.clear();
    ctx->step_key.clear();
}

namespace fg {

// Arnoldi step i from a captured hipGraph. A step's launches depend only on (n, basis and work
// vectors, the matrix and its preconditioner, the small GMRES arrays, the orthogonalization and
// preconditioner modes, the basis size): each step is captured once on the context's private
// capture stream, instantiated, and replayed on the context's stream for as long as that
// configuration holds (the device-side done flag makes the kernels of steps past convergence exit at
// once, as in the eagerly launched path). The steps contain kernel launches only.
static int graph_step32(femgpu_ctx *ctx, const femgpu_system *S, const GDev &G, int n, int i, float *wA, float *wB,
                        float *V, size_t ldv)
{
    const auto u = [](const void *p) { return static_cast<unsigned long long>(reinterpret_cast<uintptr_t>(p)); };
    const std::vector<unsigned long long> key = {
        static_cast<unsigned long long>(n), u(V), ldv, u(wA), u(wB), S->csr->serial, S->ilu->serial, u(G.H),
        u(G.flags), u(G.partials), u(G.h_status), static_cast<unsigned long long>(G.cgs), u(ctx->cgs_partials.p),
        static_cast<unsigned long long>(G.prec_mode), static_cast<unsigned long long>(G.kmax),
        static_cast<unsigned long long>(S->ilu->sweep_kernel), u(ctx->jac_partials.p)};
    if (ctx->step_key != key) {
        release_step_graphs(ctx);
        ctx->step_key = key;
        ctx->step_graphs.assign(static_cast<size_t>(G.kmax), nullptr);
    }
    float *wc = (i & 1) ? wB : wA, *wn = (i & 1) ? wA : wB;
    if (!ctx->step_graphs[i]) {
        if (!ctx->cap_stream) FG_HIP(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
        const hipStream_t saved = ctx->stream;
        FG_HIP(hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
        ctx->stream = ctx->cap_stream;
        const int rc = step_kernels32(ctx, S, G, n, i, wc, wn, V, ldv);
        ctx->stream = saved;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->cap_stream, &g);
        if (rc != FEMGPU_OK || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            set_error("mixed GMRES: capture of Arnoldi step %d failed", i);
            return rc != FEMGPU_OK ? rc : FEMGPU_ERR_HIP;
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FG_CHECK(ei == hipSuccess && x, FEMGPU_ERR_HIP, "mixed GMRES: instantiation of Arnoldi step %d failed", i);
        ctx->step_graphs[i] = x;
    }
    FG_HIP(hipGraphLaunch(ctx->step_graphs[i], ctx->stream));
    return FEMGPU_OK;
}

// The host keeps one step in flight ahead of the one it checks: step i is enqueued before step
// i - 1's done word is read, so the device never waits for the host between steps, and the step
// after the converged one drains (its kernels exit on the done flag). When the last contraction of
// the residual estimate predicts that the step in flight reaches the target (rho_{i-2}^2 / rho_{i-3}
// <= eps), the host waits for that step before enqueuing the next: the drained step's launches,
// about 20 us, are saved, and a wrong prediction costs one host round trip.
template <>
int run_cycle<float>(femgpu_ctx *ctx, const femgpu_system *S, GDev &G, int n, float *V, size_t ldv, float *wA,
                     float *wB, int kmax, int *steps_out)
{
    hipEvent_t ev[2];
    FG_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    FG_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const volatile double *hrho = reinterpret_cast<const volatile double *>(ctx->h_status + 256);
    float *wc = wA, *wn = wB;
    int rc = FEMGPU_OK, steps = 0, waited = -1;
    for (int i = 0; i < kmax - 1; ++i) {
        if (i >= 2) {   // steps up to i - 2 are known (not done), step i - 1 is in flight
            const double r1 = hrho[i - 2], r0 = i >= 3 ? hrho[i - 3] : hrho[kMapE0], eps = hrho[kMapEps];
            if (r0 > 0.0 && r1 > 0.0 && r1 * (r1 / r0) <= eps) {
                if (hipEventSynchronize(ev[(i - 1) & 1]) != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
                waited = i - 1;
                if (ctx->h_status[i - 1]) break;
            }
        }
        ctx->h_status[i] = 0;
        rc = G.graphs ? graph_step32(ctx, S, G, n, i, wA, wB, V, ldv) : step_kernels32(ctx, S, G, n, i, wc, wn, V, ldv);
        if (rc) break;
        if (hipEventRecord(ev[i & 1], ctx->stream) != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
        std::swap(wc, wn);
        steps = i + 1;
        if (i >= 1 && waited != i - 1) {
            if (hipEventSynchronize(ev[(i - 1) & 1]) != hipSuccess) { rc = FEMGPU_ERR_HIP; break; }
            if (ctx->h_status[i - 1]) break;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    *steps_out = steps;
    if (rc) set_error("mixed GMRES: launch failed");
    return rc;
}

static int apply_op64(femgpu_ctx *ctx, const femgpu_system *S, const double *x, const double *xs, double *xout,
                      double *w_out, const int *skip)
{
    if (S->csr) {
        FG_TRY(launch_spmv64(ctx, S->csr, x, xs, xout, w_out, skip));
    } else {
        FG_CHECK(xs == nullptr && xout == nullptr, FEMGPU_ERR_ARG, "EBE operator: internal scaling unsupported");
        FG_TRY(launch_ebemv(ctx, S->ebe, x, w_out));
    }
    if (S->ilu) FG_TRY(launch_sweep64(ctx, S->ilu, w_out, w_out, skip));
    else if (S->bdiag) FG_TRY(launch_blockdiag_apply(ctx, S->bdiag, w_out, w_out, skip));
    return FEMGPU_OK;
}

// EBE operators take unscaled input: scale into the basis first, then apply to the basis vector.
static int apply_step_input(femgpu_ctx *ctx, const femgpu_system *S, int n, const double *wc, const double *xs,
                            double *Vi, double *wn, const int *skip)
{
    if (S->csr) return apply_op64(ctx, S, wc, xs, Vi, wn, skip);
    k_scale_copy<double><<<grid_for(n) * 4, 256, 0, ctx->stream>>>(n, wc, xs, Vi, skip);
    FG_HIP(hipGetLastError());
    FG_TRY(launch_ebemv(ctx, S->ebe, Vi, wn));
    if (S->ilu) FG_TRY(launch_sweep64(ctx, S->ilu, wn, wn, skip));
    else if (S->bdiag) FG_TRY(launch_blockdiag_apply(ctx, S->bdiag, wn, wn, skip));
    return FEMGPU_OK;
}

static int sys_n(const femgpu_system *S) { return S->csr ? S->csr->n : S->ebe->NEQ; }

}  // namespace fg

using namespace fg;

extern "C" int femgpu_pgmres(femgpu_ctx *ctx, const femgpu_system *S, int kmax, int lmax, double tol,
                             const double *d_B, double *d_X, femgpu_solve_info *info)
{
    FG_CHECK(ctx && S && (S->csr || S->ebe) && d_B && d_X, FEMGPU_ERR_ARG, "femgpu_pgmres: bad arguments");
    FG_CHECK(!S->ilu || ilu64_ready(S->ilu), FEMGPU_ERR_STATE, "femgpu_pgmres: ILU(0) preconditioner not factored");
    const int n = sys_n(S);
    FG_CHECK(!S->ilu || S->ilu->n == n, FEMGPU_ERR_ARG, "femgpu_pgmres: preconditioner size mismatch");
    if (!S->ilu && S->bdiag) FG_TRY(blockdiag_check(S->bdiag, n));
    GDev G{};
    FG_TRY(make_gdev(ctx, kmax, G));
    size_t ldv;
    FG_TRY(ensure_ws<double>(ctx, n, kmax, ctx->gm_basis, ctx->gm_w, ldv));
    double *V = ctx->gm_basis.p, *wA = ctx->gm_w.p, *wB = wA + ldv, *wT = wB + ldv;
    const int grid = grid_for(n);
    FG_HIP(hipMemsetAsync(d_X, 0, sizeof(double) * n, ctx->stream));   // pgmres.c:39
    // first cycle: the residual is B itself (X = 0)
    FG_HIP(hipMemcpyAsync(wA, d_B, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    cycle_start<double>(ctx, G, n, wA, d_B, 0, 1, tol);
    FG_HIP(hipGetLastError());
    double normb = 0;
    FG_HIP(hipMemcpyAsync(&ctx->h_pinned[0], G.scal + 2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    normb = ctx->h_pinned[0];
    if (info) memset(info, 0, sizeof(*info));
    if (normb == 0.0) {   // zero right-hand side: X = 0 is exact (the reference divides by zero here)
        if (info) info->normb = 0.0;
        return FEMGPU_OK;
    }
    int total = 0, cycles = 0;
    double rho = normb;
    const double eps = tol * normb;
    do {
        if (cycles > 0) {
            // pgmres.c:59-80: u0 = -(M^-1 A X - B)
            if (S->csr) {
                FG_TRY(launch_spmv64(ctx, S->csr, d_X, nullptr, nullptr, wT, nullptr));
            } else {
                FG_TRY(launch_ebemv(ctx, S->ebe, d_X, wT));
            }
            if (S->ilu) FG_TRY(launch_sweep64(ctx, S->ilu, wT, wT, nullptr));
            else if (S->bdiag) FG_TRY(launch_blockdiag_apply(ctx, S->bdiag, wT, wT, nullptr));
            cycle_start<double>(ctx, G, n, wT, d_B, 1, 0, tol);
            FG_HIP(hipGetLastError());
            std::swap(wA, wT);
        }
        int steps = 0;
        if (S->csr) {
            FG_TRY(run_cycle<double>(ctx, S, G, n, V, ldv, wA, wB, kmax, &steps));
        } else {
            // EBE: same cycle with an explicit scaling pass into the basis
            hipEvent_t ev[2];
            FG_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
            FG_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
            double *wc = wA, *wn = wB;
            for (int i = 0; i < kmax - 1; ++i) {
                ctx->h_status[i] = 0;
                FG_TRY(apply_step_input(ctx, S, n, wc, G.invh + i, V + i * ldv, wn, G.flags));
                mgs_step<double>(ctx, G, n, i, wn, V, ldv);
                FG_HIP(hipGetLastError());
                FG_HIP(hipEventRecord(ev[i & 1], ctx->stream));
                std::swap(wc, wn);
                steps = i + 1;
                if (i >= 1) {
                    FG_HIP(hipEventSynchronize(ev[(i - 1) & 1]));
                    if (ctx->h_status[i - 1]) break;
                }
            }
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
        }
        k_cycle_finish<double, double><<<grid, 256, 0, ctx->stream>>>(G, n, V, ldv, d_X, 1.0, 0);
        FG_HIP(hipGetLastError());
        // device truth: steps and rho of this cycle
        FG_HIP(hipMemcpyAsync(&ctx->h_pinned[0], G.scal + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FG_HIP(hipMemcpyAsync(&ctx->h_pinned[1], G.flags + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));
        FG_TRY(check_device_error(ctx));
        rho = ctx->h_pinned[0];
        int dsteps;
        memcpy(&dsteps, &ctx->h_pinned[1], sizeof(int));
        total += dsteps;
        ++cycles;
        (void)steps;
    } while (rho > eps && cycles < lmax);
    if (info) {
        info->iterations = total;
        info->cycles = cycles;
        info->rho = rho;
        info->normb = normb;
        info->status = (rho > eps) ? 1 : 0;
    }
    return FEMGPU_OK;
}

extern "C" int femgpu_gmres_mixed(femgpu_ctx *ctx, const femgpu_system *S, const femgpu_mixed_opts *opt,
                                  double tol, const double *d_F, double *d_X, femgpu_solve_info *info)
{
    const bool csr = S && S->csr && S->ilu;
    const bool ebe = S && !S->csr && S->ebe && S->bdiag && !S->ilu;
    FG_CHECK(ctx && (csr || ebe) && opt && d_F && d_X, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed: needs a CSR operator with its ILU(0), or an EBE operator with its nodal "
             "block-diagonal preconditioner, and options");
    const int n = sys_n(S);
    if (csr) {
        FG_CHECK(opt->sweep_mode >= 0 && opt->sweep_mode <= 16 && opt->factor_sweeps >= 0 && opt->factor_sweeps <= 16,
                 FEMGPU_ERR_ARG, "femgpu_gmres_mixed: sweep_mode %d / factor_sweeps %d out of range", opt->sweep_mode,
                 opt->factor_sweeps);
        FG_CHECK(opt->factor_sweeps == 0 || opt->sweep_mode > 0, FEMGPU_ERR_ARG,
                 "femgpu_gmres_mixed: the parallel factorisation is applied with Jacobi sweeps (sweep_mode > 0)");
        FG_CHECK(opt->factor_sweeps > 0 || S->ilu->factored, FEMGPU_ERR_STATE, "femgpu_gmres_mixed: ILU(0) not factored");
    } else {
        FG_TRY(blockdiag_check(S->bdiag, n));
        FG_CHECK(S->ebe->has_values, FEMGPU_ERR_STATE, "femgpu_gmres_mixed: element matrices not set");
        FG_CHECK(opt->graphs == 0, FEMGPU_ERR_ARG, "femgpu_gmres_mixed: graph replay needs a CSR operator");
    }
    const int kin = opt->inner_kmax;
    GDev G{};
    FG_TRY(make_gdev(ctx, kin, G));
    G.prec_mode = opt->sweep_mode;
    G.cgs = opt->orthogonalization == 1;
    G.graphs = opt->graphs == 1;
    if (G.cgs) FG_TRY(ctx->cgs_partials.alloc(static_cast<size_t>(kin + 2) * kCgsBlocks));
    if (csr && opt->factor_sweeps > 0) {
        FG_TRY(launch_parilu32(ctx, S->ilu, opt->factor_sweeps));
        S->ilu->f32_factored = true;
    } else if (csr && opt->sweep_mode > 0) {
        FG_TRY(launch_factor_to_f32(ctx, S->ilu));
        S->ilu->f32_factored = true;
    }
    size_t ldv;
    FG_TRY(ensure_ws<float>(ctx, n, kin, ctx->gm_basis32, ctx->gm_w32, ldv));
    // partials of the products fused into the Jacobi preconditioner (||z||^2 at a cycle's start,
    // <V[0], w> in each Arnoldi step)
    const int nfz = prec32_dot_blocks(S, G.prec_mode);
    if (nfz > 0) FG_TRY(ctx->jac_partials.alloc(static_cast<size_t>(nfz)));
    FG_TRY(ctx->gm_r.alloc(static_cast<size_t>(n) + 64));
    float *V = ctx->gm_basis32.p, *wA = ctx->gm_w32.p, *wB = wA + ldv;
    double *r = ctx->gm_r.p;
    double *d_rn = ctx->scalars.p, *d_fn = ctx->scalars.p + 1;
    const int grid = grid_for(n);
    if (info) memset(info, 0, sizeof(*info));
    // X = 0 when it already passes: ||F|| = 0, tol >= 1, or no refinement allowed (status 1 then
    // unless ||F|| = 0), and for non-finite ||F|| (status 1)
    auto zero_answer = [&](double nf) -> int {
        FG_HIP(hipMemsetAsync(d_X, 0, sizeof(double) * n, ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));
        if (info) {
            const bool ok = nf == 0.0 || (std::isfinite(nf) && tol >= 1.0);
            info->normb = nf;
            info->rho = nf;
            info->true_rel_residual = nf == 0.0 ? 0.0 : (std::isfinite(nf) ? 1.0 : nf);
            info->status = ok ? 0 : 1;
        }
        return FEMGPU_OK;
    };
    double normF = 0.0;
    if (opt->max_refinements == 0 || tol >= 1.0) {
        FG_TRY(norm64_sync(ctx, n, d_F, d_fn, 0, normF));
        return zero_answer(normF);
    }
    // ||F|| is not waited for: the first refinement reads it on the device, and the host reads it
    // from the mapped slot after the first cycle's synchronisation (the cases above and ||F|| = 0 or
    // non-finite, which the first cycle then turns into NaN work, end with X = 0)
    volatile double *hmap_F = nullptr;
    FG_TRY(norm64_async(ctx, n, d_F, d_fn, 0, hmap_F));
    struct Event {
        hipEvent_t e = nullptr;
        ~Event() { if (e) (void)hipEventDestroy(e); }
    } evF;
    FG_HIP(hipEventCreateWithFlags(&evF.e, hipEventDisableTiming));
    FG_HIP(hipEventRecord(evF.e, ctx->stream));
    int refinements = 0, inner_total = 0;
    double rn = 0.0;
    for (;;) {
        // outer FP64 residual r = F - A X; on the first pass X = 0 (not yet written): r = F, ||r|| = ||F||
        const double *rr = d_F, *d_rcur = d_fn;
        if (refinements > 0) {
            const int nres = csr ? residual_sumsq_blocks(S->csr) : 0;
            if (nres > 0) {   // the residual kernel leaves the partials of ||r||^2
                FG_TRY(ctx->res_partials.alloc(static_cast<size_t>(nres)));
                FG_TRY(launch_residual64(ctx, S->csr, d_X, d_F, r, ctx->res_partials.p));
                FG_TRY(norm64_partials_sync(ctx, nres, ctx->res_partials.p, d_rn, 1, rn));
            } else {
                if (csr) {
                    FG_TRY(launch_residual64(ctx, S->csr, d_X, d_F, r));
                } else {
                    FG_TRY(launch_ebemv(ctx, S->ebe, d_X, r));
                    k_rsub<<<grid * 4, 256, 0, ctx->stream>>>(n, d_F, r);
                    FG_HIP(hipGetLastError());
                }
                FG_TRY(norm64_sync(ctx, n, r, d_rn, 1, rn));
            }
            rr = r;
            d_rcur = d_rn;
        }
        // a non-finite residual (a diverged preconditioner or non-finite data) ends the solve: status 1
        if (refinements > 0 && (!std::isfinite(rn) || rn <= tol * normF || refinements >= opt->max_refinements))
            break;
        // inner FP32 solve of M^-1 A d = M^-1 r / ||r||
        if (nfz > 0) {   // Jacobi on the sliced-ELL triangles: its first lower launch forms r / ||r|| in FP32
            FG_TRY(launch_jacobi_apply32_scaled(ctx, S->ilu, G.prec_mode, rr, d_rcur, wA, nullptr, ctx->jac_partials.p));
        } else {
            k_scale_to_f32<<<grid * 4, 256, 0, ctx->stream>>>(n, rr, d_rcur, wA);
            FG_HIP(hipGetLastError());
            if (csr) FG_TRY(prec32(ctx, S->ilu, G.prec_mode, wA, wA, nullptr));
            else FG_TRY(launch_blockdiag_apply32(ctx, S->bdiag, wA, wA, nullptr));
        }
        for (int cyc = 0; cyc < std::max(1, opt->inner_cycles); ++cyc) {
            if (cyc > 0) break;   // one inner cycle per refinement step
            cycle_start<float>(ctx, G, n, wA, nullptr, 0, 1, opt->inner_tol, nfz, ctx->jac_partials.p);
            FG_HIP(hipGetLastError());
            int steps = 0;
            FG_TRY(run_cycle<float>(ctx, S, G, n, V, ldv, wA, wB, kin, &steps));
            if (refinements == 0) {   // ||F|| (run_cycle has normally waited for a later step already)
                FG_HIP(hipEventSynchronize(evF.e));
                normF = *hmap_F;
                rn = normF;
                if (normF == 0.0 || !std::isfinite(normF)) return zero_answer(normF);
            }
            k_cycle_finish<float, double><<<grid, 256, 0, ctx->stream>>>(G, n, V, ldv, d_X, rn, refinements == 0);
            FG_HIP(hipGetLastError());
            // only the level-scheduled sweeps (mode 0) can raise the device error word; otherwise the
            // next residual is enqueued right behind the update, and its norm's wait covers both
            if (csr && G.prec_mode == 0) {
                FG_HIP(hipStreamSynchronize(ctx->stream));
                FG_TRY(check_device_error(ctx));
            }
            // the device's step count: the first step that reported the cycle done
            int dsteps = steps;
            for (int i = 0; i < steps; ++i)
                if (ctx->h_status[i]) {
                    dsteps = i + 1;
                    break;
                }
            inner_total += dsteps;
        }
        ++refinements;
    }
    if (info) {
        info->iterations = inner_total;
        info->inner_iterations = inner_total;
        info->cycles = refinements;
        info->normb = normF;
        info->rho = rn;
        info->true_rel_residual = rn / normF;
        info->status = (rn <= tol * normF) ? 0 : 1;
    }
    return FEMGPU_OK;
}

extern "C" int femgpu_gmres_mixed_batched(femgpu_ctx *ctx, const femgpu_system *S, int nsys, const int *row_off,
                                          const femgpu_mixed_opts *opt, double tol, const double *d_F, double *d_X,
                                          femgpu_batch_info *info)
{
    FG_CHECK(ctx && S && S->csr && S->ilu && opt && d_F && d_X && nsys >= 1 && row_off && info, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_batched: bad arguments");
    const int n = S->csr->n;
    FG_CHECK(row_off[0] == 0 && row_off[nsys] == n, FEMGPU_ERR_ARG,
             "femgpu_gmres_mixed_batched: row offsets must run from 0 to n = %d", n);
    for (int s = 0; s < nsys; ++s)
        FG_CHECK(row_off[s + 1] >= row_off[s], FEMGPU_ERR_ARG, "femgpu_gmres_mixed_batched: offsets not monotone");
    FG_CHECK(opt->sweep_mode >= 0 && opt->sweep_mode <= 16 && opt->factor_sweeps >= 0 && opt->factor_sweeps <= 16 &&
                 (opt->factor_sweeps == 0 || opt->sweep_mode > 0),
             FEMGPU_ERR_ARG, "femgpu_gmres_mixed_batched: bad preconditioner options");
    FG_CHECK(opt->factor_sweeps > 0 || S->ilu->factored, FEMGPU_ERR_STATE,
             "femgpu_gmres_mixed_batched: ILU(0) not factored");
    const int kin = opt->inner_kmax;
    memset(info, 0, sizeof(*info) * static_cast<size_t>(nsys));
    GDev G{};
    FG_TRY(make_gdev(ctx, kin, G));
    G.prec_mode = opt->sweep_mode;
    G.cgs = opt->orthogonalization == 1;
    G.graphs = opt->graphs == 1;
    if (G.cgs) FG_TRY(ctx->cgs_partials.alloc(static_cast<size_t>(kin + 2) * kCgsBlocks));
    if (opt->factor_sweeps > 0) {
        FG_TRY(launch_parilu32(ctx, S->ilu, opt->factor_sweeps));
        S->ilu->f32_factored = true;
    } else if (opt->sweep_mode > 0) {
        FG_TRY(launch_factor_to_f32(ctx, S->ilu));
        S->ilu->f32_factored = true;
    }
    size_t ldv;
    FG_TRY(ensure_ws<float>(ctx, n, kin, ctx->gm_basis32, ctx->gm_w32, ldv));
    // partials of the products fused into the Jacobi preconditioner (||z||^2 at a cycle's start,
    // <V[0], w> in each Arnoldi step)
    const int nfz = prec32_dot_blocks(S, G.prec_mode);
    if (nfz > 0) FG_TRY(ctx->jac_partials.alloc(static_cast<size_t>(nfz)));
    FG_TRY(ctx->gm_r.alloc(static_cast<size_t>(n) + 64));
    FG_TRY(ctx->bt_rowoff.upload(row_off, static_cast<size_t>(nsys) + 1, ctx->stream));
    FG_TRY(ctx->bt_sys.alloc(static_cast<size_t>(n) + 1));
    FG_TRY(ctx->bt_scal.alloc(3 * static_cast<size_t>(nsys)));
    double *d_norm = ctx->bt_scal.p, *d_scale = d_norm + nsys, *d_rn = d_scale + nsys;
    float *V = ctx->gm_basis32.p, *wA = ctx->gm_w32.p, *wB = wA + ldv;
    double *r = ctx->gm_r.p;
    const int grid = grid_for(n);
    if (n > 0) k_sys_of_row<<<grid * 4, 256, 0, ctx->stream>>>(n, nsys, ctx->bt_rowoff.p, ctx->bt_sys.p);
    FG_HIP(hipMemsetAsync(d_X, 0, sizeof(double) * n, ctx->stream));
    std::vector<double> normF(nsys), rn(nsys), scale(nsys);
    std::vector<int> refin(nsys, 0), done(nsys, 0);
    int inner_total = 0, refinements = 0;
    auto seg_norms = [&](const double *v, std::vector<double> &out) -> int {
        k_seg_norm2<<<nsys, 256, 0, ctx->stream>>>(ctx->bt_rowoff.p, v, d_norm);
        FG_HIP(hipGetLastError());
        FG_HIP(hipMemcpyAsync(out.data(), d_norm, sizeof(double) * nsys, hipMemcpyDeviceToHost, ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));
        for (double &x : out) x = sqrt(x);
        return FEMGPU_OK;
    };
    FG_TRY(seg_norms(d_F, normF));
    for (;;) {
        if (refinements == 0) {
            FG_HIP(hipMemcpyAsync(r, d_F, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            FG_TRY(launch_residual64(ctx, S->csr, d_X, d_F, r));
        }
        FG_TRY(seg_norms(r, rn));
        int active = 0;
        for (int s = 0; s < nsys; ++s) {
            if (!done[s] && (rn[s] <= tol * normF[s] || refin[s] >= opt->max_refinements)) done[s] = 1;
            scale[s] = done[s] ? 0.0 : 1.0 / rn[s];
            active += !done[s];
        }
        if (active == 0) break;
        for (int s = 0; s < nsys; ++s) refin[s] += !done[s];
        FG_HIP(hipMemcpyAsync(d_scale, scale.data(), sizeof(double) * nsys, hipMemcpyHostToDevice, ctx->stream));
        FG_HIP(hipMemcpyAsync(d_rn, rn.data(), sizeof(double) * nsys, hipMemcpyHostToDevice, ctx->stream));
        k_scale_to_f32_seg<<<grid * 4, 256, 0, ctx->stream>>>(n, ctx->bt_sys.p, r, d_scale, wA);
        FG_HIP(hipGetLastError());
        FG_TRY(prec32(ctx, S->ilu, G.prec_mode, wA, wA, nullptr, nullptr, nfz ? ctx->jac_partials.p : nullptr));
        cycle_start<float>(ctx, G, n, wA, nullptr, 0, 1, opt->inner_tol, nfz, ctx->jac_partials.p);
        FG_HIP(hipGetLastError());
        int steps = 0;
        FG_TRY(run_cycle<float>(ctx, S, G, n, V, ldv, wA, wB, kin, &steps));
        k_cycle_finish_seg<<<grid, 256, 0, ctx->stream>>>(G, n, V, ldv, d_X, ctx->bt_sys.p, d_rn);
        FG_HIP(hipGetLastError());
        FG_HIP(hipMemcpyAsync(&ctx->h_pinned[1], G.flags + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));
        FG_TRY(check_device_error(ctx));
        int dsteps;
        memcpy(&dsteps, &ctx->h_pinned[1], sizeof(int));
        inner_total += dsteps;
        ++refinements;
        for (int s = 0; s < nsys; ++s)
            if (!done[s]) info[s].inner_iterations += dsteps;
    }
    for (int s = 0; s < nsys; ++s) {
        info[s].refinements = refin[s];
        info[s].normF = normF[s];
        info[s].true_rel_residual = normF[s] > 0 ? rn[s] / normF[s] : 0.0;
        info[s].status = (rn[s] <= tol * normF[s]) ? 0 : 1;
    }
    (void)inner_total;
    return FEMGPU_OK;
}