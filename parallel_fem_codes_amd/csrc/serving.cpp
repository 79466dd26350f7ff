// Serving loop: steady transport solve requests on meshes of different sizes, solved in mixed
// precision on one GPU with continuous batching. No reference counterpart (the reference solves one
// system per MPI job).
//
//   submit   -> the request's host arrays are copied and queued (any thread)
//   prepare  -> host threads number the unknowns, build lm and the SYMRCM-ordered CSR structure
//   execute  -> executor threads, each with its own HIP stream and library context, take prepared
//               requests in arrival order: a large system runs alone, consecutive small ones are
//               packed into one block-diagonal system (one assembly per request into its slice of
//               the shared CSR values, then one batched mixed-precision solve with per-system
//               FP64 verification and convergence masks); admission keeps the device memory of
//               the work in flight under the budget, so requests that arrive while others are on
//               the GPU start as soon as an executor and memory free up, and the executors' work
//               overlaps on the device
//   wait     -> the caller collects the nodal solution and the timings
// Batches of small requests run through the block-diagonal solver, whose kernels spread the batch
// over the whole GPU, or (small_solver 1 or 2) as one persistent launch with a workgroup per system
// (femgpu_gmres_mixed_wg: each request's answer is the same alone and in any batch). A solve alone
// is faster on one compute unit below about 3000 rows (measured: 2x at 300-1000 rows, equal near
// 3000, 3-8x slower at 10000-20000), so mode 1 takes the persistent launch only when every system
// of the batch has at most kWgAutoRows rows. Units whose matrix rows exceed the sliced-ELL factor
// width, and single large requests, always use the block-diagonal solver.
// Two launch parameters are autotuned online: the batch size when max_batch = 0, then the persistent
// launch's threads per workgroup (the waves sharing a system's rows) when wg_threads = 0. The
// executors cycle through the candidates, record the throughput (unknowns solved per second) each
// achieves over two units, and keep the best.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id = 0;
    int nnodes = 0, nel = 0;
    std::vector<double> x, y, ubc;
    std::vector<int> type, ien;
    femgpu_transport_problem pb{};
    // prepared (node arrays renumbered along a Morton curve; order[new] = the request's node)
    int n = 0, nnz = 0;
    std::vector<int> invP, IA, JA, scheme, order;
    size_t bytes = 0;
    int prep_rc = FEMGPU_OK;
    std::string prep_err;
    // result
    std::vector<double> u;
    femgpu_request_info info{};
    bool done = false;
};

using Req = std::shared_ptr<Request>;

constexpr int kWgAutoRows = 3072;

// Online choice among a few candidates: sample each twice (fewest samples first), then keep the one
// with the best mean throughput.
struct Tuner {
    std::vector<int> cand;
    std::vector<double> rate;
    std::vector<int> n;
    int chosen = -1;
    explicit Tuner(std::vector<int> c) : cand(std::move(c)), rate(cand.size(), 0.0), n(cand.size(), 0) {}
    // the value to use; idx = the candidate being sampled, or -1 once settled
    int pick(int &idx)
    {
        idx = -1;
        if (chosen >= 0) return cand[chosen];
        size_t best = 0;
        for (size_t c = 1; c < cand.size(); ++c)
            if (n[c] < n[best]) best = c;
        if (n[best] >= 2) {
            size_t b = 0;
            for (size_t c = 1; c < cand.size(); ++c)
                if (rate[c] / n[c] > rate[b] / n[b]) b = c;
            chosen = static_cast<int>(b);
            return cand[b];
        }
        idx = static_cast<int>(best);
        return cand[best];
    }
    void record(int idx, double r)
    {
        if (idx < 0 || chosen >= 0) return;
        rate[idx] += r;
        n[idx] += 1;
    }
    int settled() const { return chosen >= 0 ? cand[chosen] : 0; }
};

}  // namespace

struct femgpu_server {
    int device = 0;
    femgpu_server_opts o{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_prep, cv_exec, cv_done;
    std::deque<Req> submitted;              // waiting for host preparation
    std::map<long long, Req> prepared;      // ready for the GPU, by arrival id
    std::map<long long, Req> all;           // every request not yet collected
    long long next_id = 0, next_batch = 0;
    size_t inflight = 0;                    // device bytes of the work on the GPU
    int preparing = 0;                      // requests taken by a preparation thread
    bool stop = false;
    std::vector<std::thread> threads;
    // online tuning: batch size, then threads per workgroup of the persistent launch
    Tuner batch_tuner{{4, 8, 16, 32, 64}};
    Tuner wg_tuner{{256, 512, 1024}};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

size_t footprint(const femgpu_server_opts &o, int n, int nnz, int nel)
{
    // CSR values (FP64 + FP32), columns, sliced ELL, parallel-ILU iterates and update lists,
    // Jacobi ELL, FP32 basis of kmax+1 vectors, solver vectors, mesh and element arrays
    return static_cast<size_t>(nnz) * 64 + static_cast<size_t>(n) * (4 * (o.kmax + 1) + 96) +
           static_cast<size_t>(nel) * 200 + (1u << 20);
}

// The low 21 bits of v spread to the even bit positions (one coordinate of a 2D Morton code).
uint64_t spread2(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
}

// Nodes renumbered along a Morton (Z-order) curve of their coordinates before anything else is
// built: the unknown numbering, the element pattern and the reordering then run on a numbering
// with spatial locality (a Delaunay mesh of random points arrives in random node order, which makes
// every later pass a random walk through memory). Elements keep their order; the solution is
// mapped back to the request's node order.
void spatial_renumber(Request &r)
{
    const int nn = r.nnodes;
    double x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    if (nn > 0) {
        x0 = x1 = r.x[0];
        y0 = y1 = r.y[0];
    }
    for (int v = 1; v < nn; ++v) {
        x0 = std::min(x0, r.x[v]);
        x1 = std::max(x1, r.x[v]);
        y0 = std::min(y0, r.y[v]);
        y1 = std::max(y1, r.y[v]);
    }
    const double sx = x1 > x0 ? 2097151.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 2097151.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint64_t, int>> key(nn);
    for (int v = 0; v < nn; ++v) {
        const uint64_t qx = static_cast<uint64_t>((r.x[v] - x0) * sx), qy = static_cast<uint64_t>((r.y[v] - y0) * sy);
        key[v] = {spread2(qx) | (spread2(qy) << 1), v};
    }
    std::sort(key.begin(), key.end());
    r.order.resize(nn);
    std::vector<int> newid(nn);
    for (int i = 0; i < nn; ++i) {
        r.order[i] = key[i].second;
        newid[key[i].second] = i;
    }
    std::vector<double> x(nn), y(nn), ubc(nn);
    std::vector<int> type(nn);
    for (int i = 0; i < nn; ++i) {
        const int v = r.order[i];
        x[i] = r.x[v];
        y[i] = r.y[v];
        ubc[i] = r.ubc[v];
        type[i] = r.type[v];
    }
    r.x.swap(x);
    r.y.swap(y);
    r.ubc.swap(ubc);
    r.type.swap(type);
    for (auto &v : r.ien) v = newid[v];
}

void prepare(Request &r, const femgpu_server_opts &o)
{
    for (int v : r.ien)
        if (v < 0 || v >= r.nnodes) {
            r.prep_rc = FEMGPU_ERR_ARG;
            r.prep_err = "vertex out of range";
            return;
        }
    spatial_renumber(r);
    r.invP.assign(r.nnodes, -1);
    int neq = femgpu_number_unknowns(r.nnodes, 1, r.type.data(), r.invP.data());
    if (neq < 0) {
        r.prep_rc = neq;
        r.prep_err = femgpu_last_error();
        return;
    }
    std::vector<int> lm(static_cast<size_t>(r.nel) * 3);
    int rc = femgpu_fill_lm(r.nel, 1, neq, r.ien.data(), r.type.data(), r.invP.data(), lm.data());
    std::vector<int> perm;
    const int nnz = rc == FEMGPU_OK ? fg::csr_init_vec(r.nel, 3, neq, lm.data(), 1, 0, r.nnodes, r.type.data(),
                                                       r.invP.data(), r.IA, r.JA, r.scheme, perm)
                                    : rc;
    if (nnz < 0) {
        r.prep_rc = nnz;
        r.prep_err = femgpu_last_error();
        return;
    }
    if (r.JA.empty()) r.JA.push_back(0);
    r.n = neq;
    r.nnz = nnz;
    r.bytes = footprint(o, neq, nnz, r.nel);
}

void prep_loop(femgpu_server *S)
{
    for (;;) {
        Req r;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_prep.wait(lk, [&] { return S->stop || !S->submitted.empty(); });
            if (S->stop && S->submitted.empty()) return;
            r = S->submitted.front();
            S->submitted.pop_front();
            ++S->preparing;
        }
        const double t_prep = S->now();
        prepare(*r, S->o);
        r->info.prep_s = S->now() - t_prep;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->prepared[r->id] = r;
            --S->preparing;
        }
        S->cv_exec.notify_all();
    }
}

int batch_limit(femgpu_server *S, int &cand)
{
    cand = -1;
    if (S->o.max_batch > 0) return S->o.max_batch;
    return S->batch_tuner.pick(cand);
}

// Threads per workgroup for a persistent launch: fixed, the library's size rule (0) while the
// batch size is still being tuned, then the workgroup tuner.
int wg_threads(femgpu_server *S, int &cand)
{
    cand = -1;
    if (S->o.wg_threads > 0) return S->o.wg_threads;
    if (S->o.max_batch == 0 && S->batch_tuner.chosen < 0) return 0;
    return S->wg_tuner.pick(cand);
}

// Solve one unit (one large request or a batch of small ones) on the executor's context.
// What an executor keeps across units: one matrix, a pool of meshes and the right-hand side and
// solution vectors, re-planned or grown for each unit instead of allocated and freed.
struct Executor {
    femgpu_ctx *ctx = nullptr;
    femgpu_csr *A = nullptr;
    std::vector<femgpu_mesh *> meshes;
    fg::DevBuf<double> F, X;
    ~Executor()
    {
        for (auto *M : meshes) femgpu_mesh_destroy(M);
        if (A) femgpu_csr_destroy(A);
    }
};

void run_unit(femgpu_server *S, Executor &E, std::vector<Req> &unit, long long batch_id, int threads)
{
    femgpu_ctx *ctx = E.ctx;
    const int K = static_cast<int>(unit.size());
    std::vector<int> row_off(K + 1, 0);
    std::vector<long long> nnz_off(K + 1, 0);
    for (int s = 0; s < K; ++s) {
        row_off[s + 1] = row_off[s] + unit[s]->n;
        nnz_off[s + 1] = nnz_off[s] + unit[s]->nnz;
    }
    const int N = row_off[K];
    const int NNZ = static_cast<int>(nnz_off[K]);
    std::vector<int> IA(static_cast<size_t>(N) + 1), JA(std::max(NNZ, 1));
    for (int s = 0; s < K; ++s) {
        const Request &r = *unit[s];
        for (int i = 0; i < r.n; ++i) IA[row_off[s] + i] = r.IA[i] + static_cast<int>(nnz_off[s]);
        for (int p = 0; p < r.nnz; ++p) JA[nnz_off[s] + p] = r.JA[p] + row_off[s];
    }
    IA[N] = NNZ;
    int rc = FEMGPU_OK, used_threads = 0;
    std::string err;
    femgpu_csr *A = nullptr;
    double *dF = nullptr, *dX = nullptr;
    std::vector<femgpu_batch_info> binfo(K);
    const double t_start = S->now();
    double t_setup = t_start;
    auto fail = [&](int code) {
        if (rc == FEMGPU_OK) {
            rc = code;
            err = femgpu_last_error();
        }
    };
    if (N > 0) {
        const int crc = E.A ? femgpu_csr_rebuild(E.A, N, NNZ, IA.data(), JA.data())
                            : femgpu_csr_create_ex(ctx, N, NNZ, IA.data(), JA.data(), FEMGPU_PLANS_MIXED, &E.A);
        if (crc != FEMGPU_OK) fail(FEMGPU_ERR_ARG);
        A = E.A;
        if (rc == FEMGPU_OK && (E.F.alloc(static_cast<size_t>(N) + 64) != FEMGPU_OK ||
                                E.X.alloc(static_cast<size_t>(N) + 64) != FEMGPU_OK))
            fail(FEMGPU_ERR_NOMEM);
        dF = E.F.p;
        dX = E.X.p;
        double *AA = nullptr;
        if (rc == FEMGPU_OK && femgpu_csr_values_ptr(A, &AA) != FEMGPU_OK) fail(FEMGPU_ERR_ARG);
        if (rc == FEMGPU_OK && hipMemsetAsync(dX, 0, sizeof(double) * N, ctx->stream) != hipSuccess) fail(FEMGPU_ERR_HIP);
        // one assembly per request into its slice of the shared values (ascending order: a slice's
        // dummy slot is the next slice's first entry until that slice is assembled)
        if (static_cast<int>(E.meshes.size()) < K) E.meshes.resize(K, nullptr);
        for (int s = 0; s < K && rc == FEMGPU_OK; ++s) {
            const Request &r = *unit[s];
            if (r.n == 0) continue;
            femgpu_mesh *&M = E.meshes[s];
            const int mrc = M ? femgpu_mesh_rebuild(M, r.nnodes, r.nel, r.n, r.x.data(), r.y.data(), r.type.data(),
                                                    r.invP.data(), r.ubc.data(), r.ien.data(), nullptr, r.scheme.data(),
                                                    r.nnz)
                              : femgpu_mesh_create(ctx, r.nnodes, r.nel, r.n, r.x.data(), r.y.data(), r.type.data(),
                                                   r.invP.data(), r.ubc.data(), r.ien.data(), nullptr, r.scheme.data(),
                                                   r.nnz, &M);
            if (mrc != FEMGPU_OK ||
                femgpu_assemble_transport(M, &r.pb, dX + row_off[s], nullptr, dF + row_off[s], AA + nnz_off[s]) !=
                    FEMGPU_OK)
                fail(FEMGPU_ERR_ARG);
        }
        if (rc == FEMGPU_OK && femgpu_csr_values_changed(A) != FEMGPU_OK) fail(FEMGPU_ERR_ARG);
        if (rc == FEMGPU_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) fail(FEMGPU_ERR_HIP);
        t_setup = S->now();
        femgpu_system sys = {A, nullptr, A, nullptr};
        femgpu_mixed_opts mo = S->o.mixed;
        mo.inner_kmax = S->o.kmax;
        if (rc == FEMGPU_OK) {
            const bool wg = threads >= 0 && A->sell_ok && A->par_ok;
            used_threads = wg ? (threads > 0 ? threads : fg::wg_threads_for(N, K)) : 0;
            const int st = wg ? femgpu_gmres_mixed_wg(ctx, &sys, K, row_off.data(), &mo, S->o.tol, used_threads, dF, dX,
                                                      binfo.data())
                              : femgpu_gmres_mixed_batched(ctx, &sys, K, row_off.data(), &mo, S->o.tol, dF, dX,
                                                           binfo.data());
            if (st != FEMGPU_OK) fail(FEMGPU_ERR_ARG);
        }
    }
    std::vector<double> X(static_cast<size_t>(N) + 1, 0.0);
    if (rc == FEMGPU_OK && N > 0 &&
        (hipMemcpyAsync(X.data(), dX, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        fail(FEMGPU_ERR_HIP);
    const double t_end = S->now();
    // results
    for (int s = 0; s < K; ++s) {
        Request &r = *unit[s];
        r.u.assign(r.nnodes, 0.0);
        for (int v = 0; v < r.nnodes; ++v)   // back to the request's node order
            r.u[r.order[v]] = r.type[v] == 0 ? r.ubc[v] : X[row_off[s] + r.invP[v]];
        r.info.n = r.n;
        r.info.nnz = r.nnz;
        r.info.batch_id = static_cast<int>(batch_id);
        r.info.batch_size = K;
        r.info.iterations = binfo[s].inner_iterations;
        r.info.refinements = binfo[s].refinements;
        r.info.status = rc == FEMGPU_OK ? binfo[s].status : rc;
        r.info.wg_threads = used_threads;
        r.info.true_rel_residual = binfo[s].true_rel_residual;
        r.info.t_start = t_start;
        r.info.t_end = t_end;
        r.info.setup_s = t_setup - t_start;
        r.info.solve_s = t_end - t_setup;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto &r : unit) {
        r->done = true;
        S->inflight -= r->bytes;
    }
    (void)err;
}

void exec_loop(femgpu_server *S)
{
    if (hipSetDevice(S->device) != hipSuccess) return;
    hipStream_t stream = nullptr;
    femgpu_ctx *ctx = nullptr;
    auto E = std::make_unique<Executor>();
    if (hipStreamCreate(&stream) != hipSuccess || femgpu_create(S->device, stream, &ctx) != FEMGPU_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_exec.notify_all();
        S->cv_prep.notify_all();
        return;
    }
    for (;;) {
        std::vector<Req> unit;
        long long batch_id = 0;
        int cand = -1, wcand = -1, threads = -1;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop && S->prepared.empty() && S->submitted.empty() && S->preparing == 0) break;
                if (!S->prepared.empty()) {
                    // arrival order: the oldest prepared request leads the unit
                    Req head = S->prepared.begin()->second;
                    const size_t budget = S->o.hbm_budget;
                    const bool fits_alone = S->inflight == 0 || S->inflight + head->bytes <= budget;
                    if (head->prep_rc != FEMGPU_OK || fits_alone) {
                        unit.push_back(head);
                        S->prepared.erase(S->prepared.begin());
                        if (head->prep_rc == FEMGPU_OK && head->n <= S->o.small_rows) {
                            const int limit = batch_limit(S, cand);
                            size_t bytes = head->bytes;
                            int rows = head->n;
                            while (!S->prepared.empty() && static_cast<int>(unit.size()) < limit) {
                                Req nx = S->prepared.begin()->second;
                                if (nx->prep_rc != FEMGPU_OK || nx->n > S->o.small_rows) break;
                                if (rows + nx->n > S->o.batch_rows) break;
                                if (S->inflight + bytes + nx->bytes > budget) break;
                                unit.push_back(nx);
                                S->prepared.erase(S->prepared.begin());
                                bytes += nx->bytes;
                                rows += nx->n;
                            }
                        }
                        for (auto &r : unit) S->inflight += r->prep_rc == FEMGPU_OK ? r->bytes : 0;
                        batch_id = S->next_batch++;
                        // small requests (alone or packed) may go to the persistent launch
                        int maxn = 0;
                        for (auto &r : unit) maxn = std::max(maxn, r->prep_rc == FEMGPU_OK ? r->n : 0);
                        const bool persistent = S->o.small_solver == 2 || (S->o.small_solver == 1 && maxn <= kWgAutoRows);
                        if (persistent && head->prep_rc == FEMGPU_OK && head->n <= S->o.small_rows && S->o.kmax <= 64)
                            threads = wg_threads(S, wcand);
                        break;
                    }
                }
                S->cv_exec.wait(lk);
            }
        }
        if (unit.empty()) break;
        // requests whose preparation failed are answered without the GPU
        std::vector<Req> good;
        for (auto &r : unit) {
            if (r->prep_rc != FEMGPU_OK) {
                std::lock_guard<std::mutex> lk(S->mu);
                r->info.status = r->prep_rc;
                r->info.t_start = r->info.t_end = S->now();
                r->done = true;
            } else {
                good.push_back(r);
            }
        }
        if (!good.empty()) {
            E->ctx = ctx;
            run_unit(S, *E, good, batch_id, threads);
            double rows = 0;
            for (auto &r : good) rows += r->n;
            const double dt = good[0]->info.t_end - good[0]->info.t_start;
            std::lock_guard<std::mutex> lk(S->mu);
            if (dt > 0 && good.size() > 1) S->batch_tuner.record(cand, rows / dt);
            if (dt > 0 && good.size() > 1 && good[0]->info.wg_threads > 0) S->wg_tuner.record(wcand, rows / dt);
        }
        S->cv_done.notify_all();
        S->cv_exec.notify_all();
    }
    E.reset();   // the executor's matrix and meshes before their context
    femgpu_destroy(ctx);
    (void)hipStreamDestroy(stream);
}

}  // namespace

extern "C" int femgpu_server_create(int device, const femgpu_server_opts *o, femgpu_server **out)
{
    FG_CHECK(o && out && o->executors >= 1 && o->executors <= 8 && o->prep_threads >= 1 && o->prep_threads <= 64 &&
                 o->small_rows >= 0 && o->batch_rows >= 1 && o->max_batch >= 0 && o->kmax >= 2 && o->kmax <= 150 &&
                 o->tol > 0 && o->hbm_budget > 0 && o->small_solver >= 0 && o->small_solver <= 2 &&
                 (o->wg_threads == 0 || o->wg_threads == 256 || o->wg_threads == 512 || o->wg_threads == 1024),
             FEMGPU_ERR_ARG, "femgpu_server_create: bad options");
    auto *S = new femgpu_server();
    S->device = device;
    S->o = *o;
    S->t0 = Clock::now();
    for (int i = 0; i < o->prep_threads; ++i) S->threads.emplace_back(prep_loop, S);
    for (int i = 0; i < o->executors; ++i) S->threads.emplace_back(exec_loop, S);
    *out = S;
    return FEMGPU_OK;
}

extern "C" double femgpu_server_clock(femgpu_server *S) { return S ? S->now() : 0.0; }

extern "C" int femgpu_server_submit(femgpu_server *S, int nnodes, int nel, const double *x, const double *y,
                                    const int *type, const double *ubc, const int *ien,
                                    const femgpu_transport_problem *pb, long long *ticket)
{
    FG_CHECK(S && nnodes > 0 && nel >= 0 && x && y && type && ubc && ien && pb && ticket, FEMGPU_ERR_ARG,
             "femgpu_server_submit: bad arguments");
    for (int v = 0; v < nnodes; ++v)
        FG_CHECK(type[v] == 0 || type[v] == 1, FEMGPU_ERR_ARG, "femgpu_server_submit: node %d has Type %d", v, type[v]);
    for (size_t q = 0; q < static_cast<size_t>(nel) * 3; ++q)
        FG_CHECK(ien[q] >= 0 && ien[q] < nnodes, FEMGPU_ERR_ARG, "femgpu_server_submit: vertex out of range");
    auto r = std::make_shared<Request>();
    r->nnodes = nnodes;
    r->nel = nel;
    r->x.assign(x, x + nnodes);
    r->y.assign(y, y + nnodes);
    r->ubc.assign(ubc, ubc + nnodes);
    r->type.assign(type, type + nnodes);
    r->ien.assign(ien, ien + static_cast<size_t>(nel) * 3);
    r->pb = *pb;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        FG_CHECK(!S->stop, FEMGPU_ERR_STATE, "femgpu_server_submit: server stopped");
        r->id = S->next_id++;
        r->info.t_submit = S->now();
        S->submitted.push_back(r);
        S->all[r->id] = r;
        *ticket = r->id;
    }
    S->cv_prep.notify_one();
    return FEMGPU_OK;
}

extern "C" int femgpu_server_wait(femgpu_server *S, long long ticket, double timeout_s, double *u,
                                  femgpu_request_info *info)
{
    FG_CHECK(S, FEMGPU_ERR_ARG, "femgpu_server_wait: null");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    FG_CHECK(it != S->all.end(), FEMGPU_ERR_ARG, "femgpu_server_wait: unknown ticket %lld", ticket);
    Req r = it->second;
    const bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; });
    if (!ok) return 1;
    if (u) std::copy(r->u.begin(), r->u.end(), u);
    if (info) *info = r->info;
    S->all.erase(it);
    return FEMGPU_OK;
}

extern "C" int femgpu_server_destroy(femgpu_server *S)
{
    if (!S) return FEMGPU_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_prep.notify_all();
    S->cv_exec.notify_all();
    for (auto &t : S->threads) t.join();
    delete S;
    return FEMGPU_OK;
}

extern "C" int femgpu_server_batch_tuning(femgpu_server *S, int *chosen)
{
    FG_CHECK(S && chosen, FEMGPU_ERR_ARG, "femgpu_server_batch_tuning: null");
    std::lock_guard<std::mutex> lk(S->mu);
    *chosen = S->o.max_batch > 0 ? S->o.max_batch : S->batch_tuner.settled();
    return FEMGPU_OK;
}

extern "C" int femgpu_server_wg_tuning(femgpu_server *S, int *chosen)
{
    FG_CHECK(S && chosen, FEMGPU_ERR_ARG, "femgpu_server_wg_tuning: null");
    std::lock_guard<std::mutex> lk(S->mu);
    *chosen = S->o.wg_threads > 0 ? S->o.wg_threads : S->wg_tuner.settled();
    return FEMGPU_OK;
}
