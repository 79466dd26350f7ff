// C ABI of libfemgpu.so: contexts, device-resident CSR/EBE operators, ILU(0), BLAS and PCG.
// Solvers are in krylov.hip, assembly in assembly.hip, integer structure work in structure.cpp.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace fg {

int check_device_error(femgpu_ctx *ctx)
{
    int e = 0;
    FG_HIP(hipMemcpy(&e, ctx->err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (e == 0) return FEMGPU_OK;
    FG_HIP(hipMemset(ctx->err.p, 0, sizeof(int)));
    if (e == 2) {
        set_error("ILU(0): zero pivot");
        return FEMGPU_ERR_ZERO_PIVOT;
    }
    if (e == 3) {
        set_error("sweep scheduler: ticket out of range (a previous sweep was aborted)");
        return FEMGPU_ERR_STATE;
    }
    set_error("sweep scheduler: a bounded wait expired");
    return FEMGPU_ERR_TIMEOUT;
}

__global__ void k_axpy(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] += a * x[i];
}

__global__ void k_scal(int n, double a, double *__restrict__ x)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = a * x[i];
}

static int vgrid(int n) { return std::max(1, std::min(4096, (n + 255) / 256)); }

// Host ILU(0) with the serial order of ILUp (ILUp_precond_setup.c:84-166), p = 0.
static int host_ilu0(int n, const std::vector<int> &IA, const std::vector<int> &JA, const std::vector<int> &diag,
                     std::vector<double> &LU, std::vector<double> &Dinv)
{
    std::vector<int> jw(n, -1);
    for (int i = 0; i < n; ++i) {
        for (int p = IA[i]; p < IA[i + 1]; ++p) jw[JA[p]] = p;
        double d = LU[diag[i]];
        for (int j = IA[i]; j < diag[i]; ++j) {
            const int jrow = JA[j];
            LU[j] *= Dinv[jrow];
            const double l = LU[j];
            for (int k = diag[jrow] + 1; k < IA[jrow + 1]; ++k) {
                const int col = JA[k];
                if (col == i) {
                    d -= l * LU[k];
                } else if (jw[col] >= 0) {
                    LU[jw[col]] -= l * LU[k];
                }
            }
        }
        for (int p = IA[i]; p < IA[i + 1]; ++p) jw[JA[p]] = -1;
        if (d == 0.0) {
            set_error("ILU(0): zero pivot in row %d", i);
            return FEMGPU_ERR_ZERO_PIVOT;
        }
        LU[diag[i]] = d;
        Dinv[i] = 1.0 / d;
    }
    return FEMGPU_OK;
}

}  // namespace fg

using namespace fg;

// ------------------------------------------------------------------------------- no-allocation dot
// Scratch of femgpu_ddot_noalloc: the per-block partial sums of the two-kernel dot product, one
// array per GPU in the library's code object (placed when the library is loaded on that device and
// held until the process ends), so the entry point allocates nothing. In host code `scratch` only
// names the symbol: its device address is looked up once per device, with that device current, at
// context creation, and only that address is passed to kernels. Every call on one GPU shares that
// GPU's array: calls that overlap on one GPU race on it.
static __device__ __align__(256) double scratch[kReduceBlocks];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Called with `device` current.
static int resolve_scratch(int device, double **out)
{
    FG_CHECK(device >= 0 && device < kMaxDevices, FEMGPU_ERR_ARG, "femgpu_create: device %d beyond %d", device,
             kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        FG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FG_CHECK(p, FEMGPU_ERR_HIP, "femgpu_create: no device address for the scratch array");
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return FEMGPU_OK;
}

extern "C" int femgpu_create(int device, void *stream, femgpu_ctx **out)
{
    FG_CHECK(out, FEMGPU_ERR_ARG, "femgpu_create: null output");
    int count = 0;
    FG_HIP(hipGetDeviceCount(&count));
    FG_CHECK(device >= 0 && device < count, FEMGPU_ERR_ARG, "femgpu_create: device %d of %d", device, count);
    FG_HIP(hipSetDevice(device));
    auto *ctx = new femgpu_ctx();
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(stream);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cus = prop.multiProcessorCount;
        ctx->lds_limit = prop.sharedMemPerBlock;
    }
    int rc = resolve_scratch(device, &ctx->noalloc_scratch);
    if (!rc) rc = ctx->partials.alloc(2 * kReduceBlocks);
    if (!rc) rc = ctx->scalars.alloc(16);
    if (!rc) rc = ctx->err.alloc(4);
    if (!rc && hipMemset(ctx->err.p, 0, 4 * sizeof(int)) != hipSuccess) rc = FEMGPU_ERR_HIP;
    if (!rc && hipHostMalloc((void **)&ctx->h_pinned, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        rc = FEMGPU_ERR_HIP;
    // per-step solver status written by the device: 256 ints + 256 doubles
    void *st = nullptr;
    if (!rc && hipHostMalloc(&st, 256 * sizeof(int) + 256 * sizeof(double), hipHostMallocMapped) != hipSuccess)
        rc = FEMGPU_ERR_HIP;
    if (!rc) {
        ctx->h_status = static_cast<volatile int *>(st);
        memset(st, 0, 256 * sizeof(int) + 256 * sizeof(double));
        void *dp = nullptr;
        if (hipHostGetDevicePointer(&dp, st, 0) != hipSuccess) rc = FEMGPU_ERR_HIP;
        ctx->d_status = static_cast<int *>(dp);
    }
    if (rc) {
        set_error("femgpu_create: device allocations failed");
        femgpu_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return FEMGPU_OK;
}

extern "C" int femgpu_destroy(femgpu_ctx *ctx)
{
    if (!ctx) return FEMGPU_OK;
    release_step_graphs(ctx);
    if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->h_status) (void)hipHostFree(const_cast<int *>(ctx->h_status));
    delete ctx;
    return FEMGPU_OK;
}

extern "C" int femgpu_sync(femgpu_ctx *ctx)
{
    FG_CHECK(ctx, FEMGPU_ERR_ARG, "femgpu_sync: null context");
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return check_device_error(ctx);
}

// ------------------------------------------------------------------------------- CSR

extern "C" int femgpu_csr_create(femgpu_ctx *ctx, int n, int nnz, const int *IA, const int *JA, femgpu_csr **out)
{
    return femgpu_csr_create_ex(ctx, n, nnz, IA, JA, FEMGPU_PLANS_FP64 | FEMGPU_PLANS_MIXED, out);
}

// The structure's checks and the position of every diagonal entry.
static int csr_check(int n, int nnz, const int *IA, const int *JA, std::vector<int> &diag)
{
    FG_CHECK(n >= 0 && nnz >= 0 && IA && JA, FEMGPU_ERR_ARG, "femgpu_csr_create: bad arguments");
    FG_CHECK(IA[0] == 0 && IA[n] == nnz, FEMGPU_ERR_ARG, "femgpu_csr_create: IA[0]=%d IA[n]=%d nnz=%d", IA[0], IA[n], nnz);
    diag.assign(static_cast<size_t>(n), -1);
    for (int i = 0; i < n; ++i) {
        FG_CHECK(IA[i] <= IA[i + 1], FEMGPU_ERR_ARG, "femgpu_csr_create: IA not monotone at row %d", i);
        for (int p = IA[i]; p < IA[i + 1]; ++p) {
            FG_CHECK(JA[p] >= 0 && JA[p] < n, FEMGPU_ERR_ARG, "femgpu_csr_create: column %d out of range", JA[p]);
            FG_CHECK(p == IA[i] || JA[p] > JA[p - 1], FEMGPU_ERR_ARG, "femgpu_csr_create: row %d not sorted", i);
            if (JA[p] == i) diag[i] = p;
        }
        FG_CHECK(diag[i] >= 0, FEMGPU_ERR_ARG, "femgpu_csr_create: row %d has no diagonal entry", i);
    }
    return FEMGPU_OK;
}

// Structure, plans and value buffers of A for (n, nnz, IA, JA); device buffers that are large
// enough are reused. A new serial number: anything captured for the previous structure is stale.
static int csr_setup(femgpu_csr *A, int n, int nnz, const int *IA, const int *JA, const std::vector<int> &diag)
{
    static std::atomic<unsigned long long> next_serial{1};
    femgpu_ctx *ctx = A->ctx;
    const int plans = A->plans;
    A->serial = next_serial++;
    A->factored = false;
    A->f32_factored = false;
    A->lu32_in_sell = false;
    A->n = n;
    A->nnz = nnz;
    A->hIA.assign(IA, IA + n + 1);
    A->hJA.assign(JA, JA + nnz);
    hipStream_t s = ctx->stream;
    int rc = A->IA.upload(IA, n + 1, s);
    if (!rc) rc = A->JA.upload(JA, nnz, s);
    if (!rc) rc = A->diag.upload(diag.data(), n, s);
    if (!rc) rc = A->AA.alloc(static_cast<size_t>(nnz) + 1);
    if (!rc) rc = A->AA32.alloc(static_cast<size_t>(nnz) + 1);
    if (!rc) rc = A->LU.alloc(static_cast<size_t>(nnz) + 1);
    if (!rc) rc = A->Dinv.alloc(static_cast<size_t>(n) + 1);
    if (plans & FEMGPU_PLANS_FP64) {   // exact factorisation and level-scheduled sweeps
        if (!rc) rc = build_packed_plan(ctx, n, IA, JA, diag.data(), false, A->plo);
        if (!rc) rc = build_packed_plan(ctx, n, IA, JA, diag.data(), true, A->pup);
        if (!rc) rc = build_stream_plan(ctx, n, IA, JA, diag.data(), false, A->slo);
        if (!rc) rc = build_stream_plan(ctx, n, IA, JA, diag.data(), true, A->sup);
        if (!rc) rc = build_stream_maps(ctx, A);
        if (!rc) rc = build_stream_factor(ctx, A, IA, JA, diag.data());
        if (!rc) rc = build_factor_plan(ctx, A, IA, JA, diag.data(), false, A->fplan);
    }
    if (!rc) rc = build_sell(ctx, A, IA, JA);
    if (plans & FEMGPU_PLANS_MIXED) {   // parallel FP32 ILU(0) and Jacobi iterations
        if (!rc) rc = build_factor_plan(ctx, A, IA, JA, diag.data(), true, A->fplan_nat);
        if (!rc) rc = build_jacobi_ell(ctx, A, IA, JA, diag.data());
        if (!rc) rc = build_parilu_sell(ctx, A, IA, JA, diag.data());
    }
    // a rebuilt matrix keeps FP64 sliced ELL only where its new structure allows it
    if (!rc && A->spmv_lanes == 0 && !sell64_possible(A)) rc = set_spmv_lanes(A, 8);
    if (!rc && hipMemsetAsync(A->AA.p, 0, sizeof(double) * (nnz + 1), s) != hipSuccess) rc = FEMGPU_ERR_HIP;
    if (!rc && sell64_active(A)) rc = launch_sell_values(ctx, A);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = FEMGPU_ERR_HIP;
    return rc;
}

extern "C" int femgpu_csr_create_ex(femgpu_ctx *ctx, int n, int nnz, const int *IA, const int *JA, int plans,
                                    femgpu_csr **out)
{
    FG_CHECK(ctx && out && (plans & ~3) == 0, FEMGPU_ERR_ARG, "femgpu_csr_create: bad arguments");
    std::vector<int> diag;
    FG_TRY(csr_check(n, nnz, IA, JA, diag));
    auto *A = new femgpu_csr();
    A->ctx = ctx;
    A->plans = plans;
    const int rc = csr_setup(A, n, nnz, IA, JA, diag);
    if (rc) {
        delete A;
        return rc;
    }
    *out = A;
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_rebuild(femgpu_csr *A, int n, int nnz, const int *IA, const int *JA)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_csr_rebuild: null");
    std::vector<int> diag;
    FG_TRY(csr_check(n, nnz, IA, JA, diag));
    return csr_setup(A, n, nnz, IA, JA, diag);
}

extern "C" int femgpu_csr_destroy(femgpu_csr *A)
{
    delete A;
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_set_values(femgpu_csr *A, const double *AA, int on_device)
{
    FG_CHECK(A && AA, FEMGPU_ERR_ARG, "femgpu_csr_set_values: bad arguments");
    femgpu_ctx *ctx = A->ctx;
    FG_HIP(hipMemcpyAsync(A->AA.p, AA, sizeof(double) * A->nnz,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    A->factored = false;
    return femgpu_csr_values_changed(A);
}

extern "C" int femgpu_csr_values_ptr(femgpu_csr *A, double **d_AA)
{
    FG_CHECK(A && d_AA, FEMGPU_ERR_ARG, "femgpu_csr_values_ptr: bad arguments");
    *d_AA = A->AA.p;
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_values_changed(femgpu_csr *A)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_csr_values_changed: null");
    A->factored = false;
    A->f32_factored = false;
    A->par64_factored = false;
    FG_TRY(launch_sell_values(A->ctx, A));
    // the FP32 CSR copy is read only where the sliced-ELL copy or its factor layout is missing
    if (A->sell_ok && A->par_ok) return FEMGPU_OK;
    return launch_to_f32(A->ctx, A->AA.p, A->AA32.p, A->nnz);
}

extern "C" int femgpu_csr_sweep_info(femgpu_csr *A, femgpu_sweep_info *info)
{
    FG_CHECK(A && info, FEMGPU_ERR_ARG, "femgpu_csr_sweep_info: null");
    memset(info, 0, sizeof(*info));
    const bool stream = A->sweep_kernel == 0 && A->slo.ok && A->sup.ok;
    info->kernel = stream ? 0 : 1;
    info->depth_lower = A->plo.depth;
    info->depth_upper = A->pup.depth;
    info->steps_lower = A->slo.steps;
    info->steps_upper = A->sup.steps;
    info->ring = A->slo.W;
    info->entries_per_row = A->slo.K;
    info->chunks_lower = A->plo.nchunks;
    info->rows_per_chunk = A->plo.R;
    info->lds_limit = static_cast<int>(A->ctx->lds_limit);
    info->factor_kernel = (A->factor_kernel == 0 && A->sfac.ok) ? 0 : 1;
    info->jacobi_column_bytes = A->ell_kl > 0 ? (A->ell16 ? 2 : 4) : 0;
    info->sell_column_bytes = A->sell_ok ? (A->sell16 ? 2 : 4) : 0;
    info->long_slots = std::max(A->slo.L, A->sup.L);
    info->factor_compact = info->factor_kernel == 0 && A->sfac.compact ? (A->sfac.structured ? 2 : 1) : 0;
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_set_factor_kernel(femgpu_csr *A, int kernel)
{
    FG_CHECK(A && (kernel == 0 || kernel == 1), FEMGPU_ERR_ARG, "femgpu_csr_set_factor_kernel: bad arguments");
    A->factor_kernel = kernel;
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_set_sweep_kernel(femgpu_csr *A, int kernel)
{
    FG_CHECK(A && (kernel == 0 || kernel == 1), FEMGPU_ERR_ARG, "femgpu_csr_set_sweep_kernel: bad arguments");
    A->sweep_kernel = kernel;
    return FEMGPU_OK;
}

extern "C" int femgpu_csrmv(femgpu_csr *A, const double *d_P, double *d_Q)
{
    FG_CHECK(A && d_P && d_Q && d_P != d_Q, FEMGPU_ERR_ARG, "femgpu_csrmv: bad arguments");
    return launch_spmv64(A->ctx, A, d_P, nullptr, nullptr, d_Q, nullptr);
}

extern "C" int femgpu_csr_autotune(femgpu_csr *A, femgpu_tune_info *info)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_csr_autotune: null");
    return csr_autotune(A, info);
}

extern "C" int femgpu_csr_set_lanes(femgpu_csr *A, int lanes)
{
    FG_CHECK(A && (lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16),
             FEMGPU_ERR_ARG, "femgpu_csr_set_lanes: lanes per row must be 1, 2, 4, 8 or 16 (0: sliced ELL)");
    return set_spmv_lanes(A, lanes);
}

extern "C" int femgpu_csrmv_f32(femgpu_csr *A, const float *d_P, float *d_Q)
{
    FG_CHECK(A && d_P && d_Q && d_P != d_Q, FEMGPU_ERR_ARG, "femgpu_csrmv_f32: bad arguments");
    return launch_spmv32(A->ctx, A, d_P, nullptr, nullptr, d_Q, nullptr);
}

extern "C" int femgpu_ilu0_factor(femgpu_csr *A)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_ilu0_factor: null");
    FG_CHECK(A->plans & FEMGPU_PLANS_FP64, FEMGPU_ERR_STATE, "femgpu_ilu0_factor: matrix created without FP64 plans");
    FG_TRY(launch_ilu_factor(A->ctx, A));
    FG_HIP(hipStreamSynchronize(A->ctx->stream));
    FG_TRY(check_device_error(A->ctx));
    A->factored = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_factor_host(femgpu_csr *A)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_ilu0_factor_host: null");
    FG_CHECK(A->plans & FEMGPU_PLANS_FP64, FEMGPU_ERR_STATE, "femgpu_ilu0_factor_host: matrix created without FP64 plans");
    femgpu_ctx *ctx = A->ctx;
    std::vector<double> LU(A->nnz), Dinv(A->n);
    std::vector<int> diag(A->n);
    FG_HIP(hipMemcpyAsync(LU.data(), A->AA.p, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipMemcpyAsync(diag.data(), A->diag.p, sizeof(int) * A->n, hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    FG_TRY(host_ilu0(A->n, A->hIA, A->hJA, diag, LU, Dinv));
    FG_HIP(hipMemcpyAsync(A->LU.p, LU.data(), sizeof(double) * A->nnz, hipMemcpyHostToDevice, ctx->stream));
    FG_HIP(hipMemcpyAsync(A->Dinv.p, Dinv.data(), sizeof(double) * A->n, hipMemcpyHostToDevice, ctx->stream));
    FG_TRY(packed_refresh(ctx, A));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    A->factored = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_factor_parallel_f32(femgpu_csr *A, int sweeps)
{
    FG_CHECK(A && sweeps >= 0 && sweeps <= 1024, FEMGPU_ERR_ARG, "femgpu_ilu0_factor_parallel_f32: bad arguments");
    FG_TRY(launch_parilu32(A->ctx, A, sweeps));
    FG_HIP(hipStreamSynchronize(A->ctx->stream));
    A->f32_factored = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_factor_to_f32(femgpu_csr *A)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_ilu0_factor_to_f32: null");
    FG_CHECK(A->factored, FEMGPU_ERR_STATE, "femgpu_ilu0_factor_to_f32: not factored");
    FG_TRY(launch_factor_to_f32(A->ctx, A));
    A->f32_factored = true;
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_apply_jacobi_f32(femgpu_csr *A, int iterations, const float *d_p, float *d_z)
{
    FG_CHECK(A && d_p && d_z, FEMGPU_ERR_ARG, "femgpu_ilu0_apply_jacobi_f32: bad arguments");
    FG_CHECK(A->f32_factored, FEMGPU_ERR_STATE, "femgpu_ilu0_apply_jacobi_f32: no FP32 factors");
    return launch_jacobi_apply32(A->ctx, A, iterations, d_p, d_z, nullptr);
}

extern "C" int femgpu_ilu0_jacobi_iteration_f32(femgpu_csr *A, int upper, const float *d_rhs, const float *d_in,
                                                float *d_out)
{
    FG_CHECK(A && d_rhs && d_out && (upper == 0 || upper == 1) && d_out != d_in && d_out != d_rhs, FEMGPU_ERR_ARG,
             "femgpu_ilu0_jacobi_iteration_f32: bad arguments");
    FG_CHECK(A->f32_factored, FEMGPU_ERR_STATE, "femgpu_ilu0_jacobi_iteration_f32: no FP32 factors");
    if (!A->ell_kl && A->lu32_in_sell) FG_TRY(launch_parilu_to_csr(A->ctx, A));
    return launch_jacobi_iteration32(A->ctx, A, upper, d_rhs, d_in, d_out);
}

extern "C" int femgpu_ilu0_get_f32(femgpu_csr *A, float *LU, float *Dinv)
{
    FG_CHECK(A && LU && Dinv, FEMGPU_ERR_ARG, "femgpu_ilu0_get_f32: bad arguments");
    FG_CHECK(A->f32_factored, FEMGPU_ERR_STATE, "femgpu_ilu0_get_f32: no FP32 factors");
    if (A->lu32_in_sell) FG_TRY(launch_parilu_to_csr(A->ctx, A));
    FG_HIP(hipMemcpyAsync(LU, A->LU32.p, sizeof(float) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
    FG_HIP(hipMemcpyAsync(Dinv, A->Dinv32.p, sizeof(float) * A->n, hipMemcpyDeviceToHost, A->ctx->stream));
    FG_HIP(hipStreamSynchronize(A->ctx->stream));
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_apply(femgpu_csr *A, const double *d_p, double *d_z)
{
    FG_CHECK(A && d_p && d_z, FEMGPU_ERR_ARG, "femgpu_ilu0_apply: bad arguments");
    FG_CHECK(ilu64_ready(A), FEMGPU_ERR_STATE, "femgpu_ilu0_apply: not factored");
    return launch_sweep64(A->ctx, A, d_p, d_z, nullptr);
}

extern "C" int femgpu_ilu0_factor_parallel(femgpu_csr *A, int sweeps)
{
    FG_CHECK(A, FEMGPU_ERR_ARG, "femgpu_ilu0_factor_parallel: null");
    return launch_parilu64(A->ctx, A, sweeps);
}

extern "C" int femgpu_ilu0_set_apply(femgpu_csr *A, int jacobi_iterations)
{
    FG_CHECK(A && jacobi_iterations >= 0 && jacobi_iterations <= 1024, FEMGPU_ERR_ARG, "femgpu_ilu0_set_apply: bad arguments");
    FG_CHECK(jacobi_iterations == 0 || A->ell_kl > 0, FEMGPU_ERR_ARG,
             "femgpu_ilu0_set_apply: Jacobi iterations need the sliced-ELL triangles (at most 16 entries each)");
    A->prec64_jacobi = jacobi_iterations;
    return FEMGPU_OK;
}

extern "C" int femgpu_ilu0_apply_f32(femgpu_csr *A, const float *d_p, float *d_z)
{
    FG_CHECK(A && d_p && d_z, FEMGPU_ERR_ARG, "femgpu_ilu0_apply_f32: bad arguments");
    FG_CHECK(A->factored, FEMGPU_ERR_STATE, "femgpu_ilu0_apply_f32: not factored");
    return launch_sweep32(A->ctx, A, d_p, d_z, nullptr);
}

extern "C" int femgpu_ilu0_get(femgpu_csr *A, double *LU, double *Dinv)
{
    FG_CHECK(A && LU && Dinv, FEMGPU_ERR_ARG, "femgpu_ilu0_get: bad arguments");
    FG_CHECK(A->factored, FEMGPU_ERR_STATE, "femgpu_ilu0_get: not factored");
    FG_HIP(hipMemcpyAsync(LU, A->LU.p, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
    FG_HIP(hipMemcpyAsync(Dinv, A->Dinv.p, sizeof(double) * A->n, hipMemcpyDeviceToHost, A->ctx->stream));
    FG_HIP(hipStreamSynchronize(A->ctx->stream));
    return FEMGPU_OK;
}

// ------------------------------------------------------------------------------- BLAS

extern "C" int femgpu_ddot(femgpu_ctx *ctx, int n, const double *x, const double *y, double *result)
{
    FG_CHECK(ctx && n >= 0 && x && y && result, FEMGPU_ERR_ARG, "femgpu_ddot: bad arguments");
    FG_TRY(dot64(ctx, n, x, y, ctx->scalars.p));
    FG_HIP(hipMemcpyAsync(result, ctx->scalars.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FEMGPU_OK;
}

extern "C" int femgpu_ddot_noalloc(femgpu_ctx *ctx, int n, const double *d_x, const double *d_y, double *d_result)
{
    FG_CHECK(ctx && ctx->noalloc_scratch && d_x && d_y && d_result, FEMGPU_ERR_ARG,
             "femgpu_ddot_noalloc: bad arguments");
    FG_CHECK(n >= 0, FEMGPU_ERR_ARG, "femgpu_ddot_noalloc: n = %d", n);
    const int grid = std::max(1, std::min(kReduceBlocks, (n + 1023) / 1024));
    FG_CHECK(grid <= kReduceBlocks, FEMGPU_ERR_ARG, "femgpu_ddot_noalloc: %d partial sums exceed the %d-entry scratch",
             grid, kReduceBlocks);
    return dot64_into(ctx, n, d_x, d_y, ctx->noalloc_scratch, d_result);
}

extern "C" int femgpu_daxpy(femgpu_ctx *ctx, int n, double a, const double *x, double *y)
{
    FG_CHECK(ctx && n >= 0 && x && y, FEMGPU_ERR_ARG, "femgpu_daxpy: bad arguments");
    if (n) k_axpy<<<vgrid(n), 256, 0, ctx->stream>>>(n, a, x, y);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

extern "C" int femgpu_dscal(femgpu_ctx *ctx, int n, double a, double *x)
{
    FG_CHECK(ctx && n >= 0 && x, FEMGPU_ERR_ARG, "femgpu_dscal: bad arguments");
    if (n) k_scal<<<vgrid(n), 256, 0, ctx->stream>>>(n, a, x);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

extern "C" int femgpu_dcopy(femgpu_ctx *ctx, int n, const double *x, double *y)
{
    FG_CHECK(ctx && n >= 0 && x && y, FEMGPU_ERR_ARG, "femgpu_dcopy: bad arguments");
    if (n) FG_HIP(hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    return FEMGPU_OK;
}

extern "C" int femgpu_dzero(femgpu_ctx *ctx, int n, double *x)
{
    FG_CHECK(ctx && n >= 0 && x, FEMGPU_ERR_ARG, "femgpu_dzero: bad arguments");
    if (n) k_scal<<<vgrid(n), 256, 0, ctx->stream>>>(n, 0.0, x);
    FG_HIP(hipGetLastError());
    return FEMGPU_OK;
}

// ------------------------------------------------------------------------------- PCG

extern "C" int femgpu_pcg(femgpu_ctx *ctx, const femgpu_system *S, int itmax, double tol, const double *d_B,
                          double *d_X, femgpu_solve_info *info)
{
    FG_CHECK(ctx && S && S->csr && d_B && d_X, FEMGPU_ERR_ARG, "femgpu_pcg: needs a CSR operator");
    FG_CHECK(!S->ilu || ilu64_ready(S->ilu), FEMGPU_ERR_STATE, "femgpu_pcg: ILU(0) not factored");
    const int n = S->csr->n;
    if (!S->ilu && S->bdiag) FG_TRY(blockdiag_check(S->bdiag, n));
    DevBuf<double> P, Q, R, Z;
    FG_TRY(P.alloc(n));
    FG_TRY(Q.alloc(n));
    FG_TRY(R.alloc(n));
    FG_TRY(Z.alloc(n));
    auto dot = [&](const double *a, const double *b, double &out) -> int {
        FG_TRY(dot64(ctx, n, a, b, ctx->scalars.p));
        FG_HIP(hipMemcpyAsync(&out, ctx->scalars.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FG_HIP(hipStreamSynchronize(ctx->stream));
        return FEMGPU_OK;
    };
    auto precond = [&](double *z) -> int {
        if (S->ilu) FG_TRY(launch_sweep64(ctx, S->ilu, z, z, nullptr));
        else if (S->bdiag) FG_TRY(launch_blockdiag_apply(ctx, S->bdiag, z, z, nullptr));
        return FEMGPU_OK;
    };
    // pcg.c:31-47
    FG_TRY(femgpu_dzero(ctx, n, d_X));
    FG_TRY(femgpu_dcopy(ctx, n, d_B, R.p));
    FG_TRY(femgpu_dcopy(ctx, n, R.p, Z.p));
    FG_TRY(precond(Z.p));
    FG_TRY(femgpu_dcopy(ctx, n, Z.p, P.p));
    double gnew = 0;
    FG_TRY(dot(R.p, Z.p, gnew));
    const double r0 = gnew;
    int iter = 0;
    if (!(gnew < tol * tol * r0)) {
        for (iter = 0; iter < itmax; ++iter) {
            // pcg.c:49-80
            FG_TRY(launch_spmv64(ctx, S->csr, P.p, nullptr, nullptr, Q.p, nullptr));
            double tau = 0;
            FG_TRY(dot(P.p, Q.p, tau));
            const double alpha = gnew / tau;
            FG_TRY(femgpu_daxpy(ctx, n, alpha, P.p, d_X));
            FG_TRY(femgpu_daxpy(ctx, n, -alpha, Q.p, R.p));
            FG_TRY(femgpu_dcopy(ctx, n, R.p, Z.p));
            FG_TRY(precond(Z.p));
            const double gold = gnew;
            FG_TRY(dot(R.p, Z.p, gnew));
            if (gnew < tol * tol * r0) break;
            FG_TRY(femgpu_dscal(ctx, n, gnew / gold, P.p));
            FG_TRY(femgpu_daxpy(ctx, n, 1.0, Z.p, P.p));
        }
    }
    FG_HIP(hipStreamSynchronize(ctx->stream));
    FG_TRY(check_device_error(ctx));
    if (info) {
        memset(info, 0, sizeof(*info));
        info->iterations = iter;
        info->rho = std::sqrt(std::fabs(gnew));
        info->normb = std::sqrt(std::fabs(r0));
        info->status = (gnew < tol * tol * r0) ? 0 : 1;
    }
    return FEMGPU_OK;
}
