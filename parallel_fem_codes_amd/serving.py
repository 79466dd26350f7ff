"""Serving workload and harness around the library's request loop (femgpu_server_*).

A request is a steady advection-diffusion solve (the operator of the steady transport driver,
01_SS_Transport_Equation) on an unstructured Delaunay triangulation of the unit square from N
uniform random points plus boundary nodes, with diffusivity k and a uniform unit velocity of random
angle; boundary conditions as the structured workload: u = 1 on x = 0, 0.25 < y < 0.5, u = 0 on the
rest of the boundary. Requests arrive as a Poisson process.

    python -m parallel_fem_codes_amd.serving --requests 200 --rate 200 --nmin 2000 --nmax 50000

prints one JSON line: p50/p99 latency and solves/s of the server (continuous batching, mixed
precision) and of a sequential plain-FP64 GPU loop over the same requests (ILU(0) + pgmres, the
reference's solver).
"""
import argparse
import ctypes as C
import json
import sys
import threading
import time

import numpy as np

from . import fem
from . import native as N


# ----------------------------------------------------------------------------- workload

def square_mesh(npts, rng):
    """Delaunay triangulation (the library's femgpu_delaunay) of npts uniform random points in the
    unit square plus boundary nodes (about sqrt(npts) per side, uniformly spaced); triangles
    counter-clockwise."""
    m = max(4, int(np.sqrt(npts)))
    t = np.linspace(0.0, 1.0, m + 1)[:-1]
    bx = np.concatenate([t, np.ones(m), 1.0 - t, np.zeros(m)])
    by = np.concatenate([np.zeros(m), t, np.ones(m), 1.0 - t])
    px = rng.uniform(0.0, 1.0, npts) * (1 - 2e-6) + 1e-6
    py = rng.uniform(0.0, 1.0, npts) * (1 - 2e-6) + 1e-6
    x = np.concatenate([bx, px])
    y = np.concatenate([by, py])
    tri, _ = fem.delaunay(x, y)
    nb = len(bx)
    typ = np.ones(len(x), np.int32)
    typ[:nb] = 0
    ubc = np.where((np.abs(x) <= 1e-15) & (y > 0.25) & (y < 0.5), 1.0, 0.0)
    return x, y, typ, ubc, tri.ravel().astype(np.int32)


def request_params(count, seed, nmin, nmax, rate, kmin=1e-5, kmax=1e-2):
    """count request descriptions: arrival time (a Poisson process of the given rate, per second),
    N and k log-uniform, the velocity angle uniform, and the seed of the request's points."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    ps = []
    for i in range(count):
        npts = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        k = float(np.exp(rng.uniform(np.log(kmin), np.log(kmax))))
        th = float(rng.uniform(0.0, 2 * np.pi))
        ps.append(dict(index=i, arrival=float(arrivals[i]), npts=npts, k=k, theta=th,
                       mesh_seed=int(rng.integers(1 << 62))))
    return ps


def build_request(p):
    """The request's mesh (Delaunay of its seeded points) and problem."""
    x, y, typ, ubc, ien = square_mesh(p["npts"], np.random.default_rng(p["mesh_seed"]))
    th = p["theta"]
    pb = fem.make_problem(k=p["k"], bx=(np.cos(th), 0.0, 0.0), by=(np.sin(th), 0.0, 0.0))
    return dict(p, x=x, y=y, type=typ, ubc=ubc, ien=ien, pb=pb)


def make_requests(count, seed, nmin, nmax, rate, kmin=1e-5, kmax=1e-2, workers=8):
    """count requests (request_params, meshes built by a thread pool)."""
    from concurrent.futures import ThreadPoolExecutor
    ps = request_params(count, seed, nmin, nmax, rate, kmin, kmax)
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(build_request, ps))


# ----------------------------------------------------------------------------- server

def server_opts(hbm_budget=64 << 30, executors=4, prep_threads=8, small_rows=20000, batch_rows=400000,
                max_batch=0, kmax=30, tol=1e-8, sweep_mode=2, factor_sweeps=2, inner_tol=1e-4, max_refinements=50,
                orthogonalization="CGS", small_solver="shared", wg_threads=0):
    """small_solver for batches of small requests: "shared" (block-diagonal shared-Krylov solver),
    "persistent" (one launch, a workgroup per system) or "auto" (persistent when every system has at
    most 3072 rows); wg_threads 0 = tuned online. Classical Gram-Schmidt by default: the serving
    systems' inner cycles run long (up to kmax - 1 steps), where MGS costs i + 3 launches per step
    and CGS four."""
    o = N.ServerOpts()
    o.hbm_budget, o.executors, o.prep_threads = int(hbm_budget), executors, prep_threads
    o.small_rows, o.batch_rows, o.max_batch, o.kmax, o.tol = small_rows, batch_rows, max_batch, kmax, tol
    o.small_solver = {"shared": 0, "auto": 1, "persistent": 2}[small_solver]
    o.wg_threads = wg_threads
    o.mixed = fem.mixed_opts(kmax, inner_tol, max_refinements, sweep_mode, factor_sweeps, orthogonalization)
    return o


class Server:
    """The library's request loop (continuous batching, mixed precision) on one GPU."""

    def __init__(self, device=0, opts=None):
        self.opts = opts if opts is not None else server_opts()
        h = C.c_void_p()
        N.check(N.lib().femgpu_server_create(int(device), C.byref(self.opts), C.byref(h)))
        self.h = h

    def clock(self):
        return N.lib().femgpu_server_clock(self.h)

    def submit(self, r):
        t = C.c_longlong()
        N.check(N.lib().femgpu_server_submit(self.h, len(r["x"]), len(r["ien"]) // 3, np.ascontiguousarray(r["x"]),
                                             np.ascontiguousarray(r["y"]), np.ascontiguousarray(r["type"]),
                                             np.ascontiguousarray(r["ubc"]), np.ascontiguousarray(r["ien"]),
                                             C.byref(r["pb"]), C.byref(t)))
        return t.value

    def wait(self, ticket, nnodes, timeout=300.0):
        u = np.empty(nnodes)
        info = N.RequestInfo()
        rc = N.lib().femgpu_server_wait(self.h, C.c_longlong(ticket), float(timeout), u.ctypes.data_as(C.c_void_p),
                                        C.byref(info))
        if rc == 1:
            raise TimeoutError(f"request {ticket} not done within {timeout} s")
        N.check(rc)
        return u, info

    def batch_size(self):
        c = C.c_int()
        N.check(N.lib().femgpu_server_batch_tuning(self.h, C.byref(c)))
        return c.value

    def wg_threads(self):
        c = C.c_int()
        N.check(N.lib().femgpu_server_wg_tuning(self.h, C.byref(c)))
        return c.value

    def close(self):
        if getattr(self, "h", None):
            N.lib().femgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_server(reqs, opts=None, device=0, warmup=0):
    """Feeds the requests at their arrival times (relative to the first), collects every result.
    warmup > 0 first solves that many of the requests one by one (results dropped), so that the
    timed stream does not include the first launches of each kernel. Returns (solutions, infos,
    wall seconds, batch size chosen, workgroup threads chosen)."""
    srv = Server(device, opts)
    for r in reqs[:warmup]:
        srv.wait(srv.submit(r), len(r["x"]))
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def feed():
        for i, r in enumerate(reqs):
            delay = r["arrival"] - reqs[0]["arrival"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(r)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    sols, infos = [], []
    for i, r in enumerate(reqs):
        u, info = srv.wait(tickets[i], len(r["x"]))
        sols.append(u)
        infos.append(info)
    wall = time.perf_counter() - t0
    batch, wg = srv.batch_size(), srv.wg_threads()
    srv.close()
    return sols, infos, wall, batch, wg


def run_stream(params, opts=None, device=0, workers=6, lookahead=48, keep=(), warmup=8, progress=30.0):
    """The serving workload as a stream: request meshes are built by a pool of client threads ahead
    of their arrival (at most `lookahead` in flight) and submitted at their arrival times; results
    are collected as they finish. Only the requests whose index is in `keep` are retained with
    their solutions (for comparisons); of the others only the server's timings remain. Returns
    (per-request dicts in arrival order, kept {index: (request, solution)}, wall seconds, client
    submit lag per request (seconds behind the arrival time), batch size, workgroup threads). A line
    on stderr reports the requests done every `progress` seconds (0: none)."""
    import queue
    from collections import deque
    from concurrent.futures import ThreadPoolExecutor
    srv = Server(device, opts)
    for p in params[:warmup]:
        r = build_request(p)
        srv.wait(srv.submit(r), len(r["x"]))
    keep = set(keep)
    count = len(params)
    lag = np.zeros(count)
    q = queue.Queue()
    kept = {}
    t0 = time.perf_counter()
    a0 = params[0]["arrival"]

    def feed():
        with ThreadPoolExecutor(workers) as ex:
            fut, nxt = deque(), 0
            for i, p in enumerate(params):
                while nxt < count and nxt < i + lookahead:
                    fut.append(ex.submit(build_request, params[nxt]))
                    nxt += 1
                r = fut.popleft().result()
                delay = p["arrival"] - a0 - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                lag[i] = max(0.0, -delay)
                q.put((i, srv.submit(r), len(r["x"]), r if i in keep else None))

    th = threading.Thread(target=feed)
    th.start()
    out = [None] * count
    t_note = t0
    for done in range(count):
        i, ticket, nn, r = q.get()
        u, info = srv.wait(ticket, nn, timeout=3600.0)
        now = time.perf_counter()
        if progress and now - t_note >= progress:
            print(f"# {done + 1}/{count} requests done after {now - t0:.1f} s", file=sys.stderr, flush=True)
            t_note = now
        out[i] = dict(n=info.n, nnz=info.nnz, latency=info.t_end - info.t_submit, prep=info.prep_s,
                      setup=info.setup_s, solve=info.solve_s, queue=info.t_start - info.t_submit - info.prep_s,
                      status=info.status, residual=info.true_rel_residual, batch=info.batch_size,
                      wg=info.wg_threads, iterations=info.iterations)
        if r is not None:
            kept[i] = (r, u)
    th.join()
    wall = time.perf_counter() - t0
    batch, wg = srv.batch_size(), srv.wg_threads()
    srv.close()
    return out, kept, wall, lag, batch, wg


def run_isolated(reqs, opts=None, device=0):
    """Each request alone on an idle server (submitted after the previous one finished): the
    isolated-solve latencies the loaded server's are compared with."""
    o = opts if opts is not None else server_opts()
    o.executors, o.max_batch = 1, 1
    srv = Server(device, o)
    lat = []
    for r in reqs:
        _, info = srv.wait(srv.submit(r), len(r["x"]))
        lat.append(info.t_end - info.t_submit)
    srv.close()
    return np.array(lat)


def solve_fp64(dev, r, kmax=30, tol=1e-8):
    """The reference's path on the GPU for one request: structures, device assembly
    (Build_K_F_SUPG), ILUp_precond_setup (factor + F := M^-1 F) and pgmres, FP64."""
    import torch
    x, y, t, ubc, ien = r["x"], r["y"], r["type"], r["ubc"], r["ien"]
    invP, neq = fem.number_unknowns(t)
    lm = fem.Fill_LM(ien, t, invP, neq)
    c = fem.csr_Initialization(lm, 3, neq, "SYMRCM", type_=t, invP_id=invP)
    M = fem.Mesh(dev, x, y, t, c.invP_id, ubc, ien, neq, scheme=c.scheme, nnz=c.nnz)
    A = fem.CsrMatrix(dev, c.IA, c.JA)
    u0 = torch.zeros(neq, dtype=torch.float64, device="cuda")
    F = torch.empty(neq, dtype=torch.float64, device="cuda")
    M.Build_K_F_SUPG(r["pb"], u0, F, AA_ptr=A.values_ptr())
    A.values_changed()
    A.ILUp_precond_setup(F)
    X = torch.empty_like(F)
    info = fem.pgmres(dev, A, F, X, kmax, 1000, tol, precond=A)
    Xh = X.cpu().numpy()
    u = np.where(t == 0, ubc, Xh[np.maximum(c.invP_id, 0)])
    return u, info


def run_sequential_fp64(reqs):
    """A sequential loop over the same requests with the FP64 path; latency under the same
    arrivals = queueing behind the previous request + its own service time."""
    import torch
    dev = fem.Device(0)
    service, sols = [], []
    for r in reqs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, info = solve_fp64(dev, r)
        torch.cuda.synchronize()
        service.append(time.perf_counter() - t0)
        sols.append(u)
    end, lat = 0.0, []
    for r, s in zip(reqs, service):
        a = r["arrival"] - reqs[0]["arrival"]
        start = max(a, end)
        end = start + s
        lat.append(end - a)
    dev.close()
    return sols, np.array(service), np.array(lat), end


def stream_main(a):
    """The stated serving workload: a long stream of requests, N log-uniform in [nmin, nmax],
    summarised into one JSON object (latency percentiles, throughput, per-stage host and device
    times, convergence), with the isolated-solve latencies and the sequential plain-FP64 GPU loop
    measured on a sample of the same requests."""
    import torch
    torch.cuda.init()
    params = request_params(a.requests, a.seed, a.nmin, a.nmax, a.rate)
    rng = np.random.default_rng(a.seed + 1)
    sample = sorted(rng.choice(len(params), size=min(a.sample, len(params)), replace=False).tolist())
    extra = {"orthogonalization": a.orth} if a.orth else {}
    opts = server_opts(executors=a.executors, prep_threads=a.prep_threads, max_batch=a.max_batch,
                       small_solver=a.small_solver, wg_threads=a.wg_threads, **extra)
    t0 = time.perf_counter()
    out, kept, wall, lag, batch, wg = run_stream(params, opts, workers=a.client_threads, keep=sample,
                                                 warmup=a.warmup)
    print(f"# stream of {len(params)} done in {wall:.1f} s", file=sys.stderr, flush=True)
    lat = np.array([o["latency"] for o in out])
    res = np.array([o["residual"] for o in out])
    ns = np.array([o["n"] for o in out])
    pct = lambda v, q: round(1e3 * float(np.percentile(v, q)), 3)
    summary = {
        "workload": f"{len(params)} steady SUPG transport solves on Delaunay meshes of N log-uniform in "
                    f"[{a.nmin}, {a.nmax}] random points plus boundary nodes, k log-uniform in [1e-5, 1e-2], "
                    f"random flow angle; Poisson arrivals at {a.rate}/s (seed {a.seed}); mixed precision, "
                    f"FP64-verified 1e-8",
        "requests": len(params), "arrival_rate_per_s": a.rate,
        "unknowns": {"min": int(ns.min()), "median": int(np.median(ns)), "max": int(ns.max()),
                     "total": int(ns.sum())},
        "server": {"solves_per_s": round(len(params) / wall, 3), "unknowns_per_s": round(float(ns.sum()) / wall, 1),
                   "wall_s": round(wall, 2), "p50_ms": pct(lat, 50), "p90_ms": pct(lat, 90), "p99_ms": pct(lat, 99),
                   "converged": int(sum(o["status"] == 0 for o in out)), "max_true_rel_residual": float(res.max()),
                   "batch_size": batch, "wg_threads": wg,
                   "requests_in_batches": int(sum(o["batch"] > 1 for o in out)),
                   "persistent_launch_requests": int(sum(o["wg"] > 0 for o in out)),
                   "mean_prep_ms": round(1e3 * float(np.mean([o["prep"] for o in out])), 3),
                   "mean_setup_ms": round(1e3 * float(np.mean([o["setup"] for o in out])), 3),
                   "mean_solve_ms": round(1e3 * float(np.mean([o["solve"] for o in out])), 3),
                   "mean_queue_ms": round(1e3 * float(np.mean([o["queue"] for o in out])), 3),
                   "client_submit_lag_ms": {"p50": pct(lag, 50), "p99": pct(lag, 99)}},
        "executors": a.executors, "prep_threads": a.prep_threads, "client_threads": a.client_threads}
    print(f"# server: {summary['server']}", file=sys.stderr, flush=True)
    reqs = [kept[i][0] for i in sample]
    if a.isolated:
        iso = run_isolated(reqs, server_opts(small_solver=a.small_solver, wg_threads=a.wg_threads,
                                             prep_threads=a.prep_threads, **extra))
        lat_s = np.array([out[i]["latency"] for i in sample])
        summary["sample"] = {"requests": len(sample), "loaded_p50_ms": pct(lat_s, 50), "loaded_p99_ms": pct(lat_s, 99),
                             "isolated_p50_ms": pct(iso, 50), "isolated_p99_ms": pct(iso, 99),
                             "p99_loaded_over_isolated": round(float(np.percentile(lat_s, 99) / np.percentile(iso, 99)), 3),
                             "median_per_request_ratio": round(float(np.median(lat_s / iso)), 3)}
        print(f"# sample: {summary['sample']}", file=sys.stderr, flush=True)
    if not a.no_fp64:
        s64, service, lat64, end = run_sequential_fp64(reqs)
        diff = max(float(np.linalg.norm(kept[i][1] - v) / max(np.linalg.norm(v), 1e-300)) for i, v in zip(sample, s64))
        ns_s = np.array([len(r["x"]) for r in reqs])
        summary["fp64_sequential_sample"] = {
            "requests": len(reqs), "mean_service_ms": round(1e3 * float(service.mean()), 3),
            "solves_per_s_capacity": round(1.0 / float(service.mean()), 3),
            "note": "the reference's FP64 path (ILU(0) + pgmres) one request at a time on the same GPU, on the "
                    "sample; its capacity is 1 / mean service time (the stream's mean N matches the sample's)",
            "sample_mean_nodes": int(ns_s.mean())}
        summary["max_rel_diff_vs_fp64"] = diff
    print(json.dumps(summary))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(summary, fh, indent=1)
            fh.write("\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=200.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--nmin", type=int, default=2000)
    ap.add_argument("--nmax", type=int, default=50000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--executors", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=0)
    ap.add_argument("--small-solver", choices=("shared", "auto", "persistent"), default="shared")
    ap.add_argument("--orth", choices=("MGS", "CGS"), default=None, help="inner Gram-Schmidt (default: server_opts')")
    ap.add_argument("--wg-threads", type=int, default=0)
    ap.add_argument("--isolated", action="store_true", help="also time every request alone on an idle server")
    ap.add_argument("--warmup", type=int, default=8, help="requests solved before the timed stream")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--stream", action="store_true",
                    help="streamed run (meshes built by client threads as requests arrive), JSON summary only")
    ap.add_argument("--prep-threads", type=int, default=8)
    ap.add_argument("--client-threads", type=int, default=6)
    ap.add_argument("--sample", type=int, default=100, help="stream: requests kept for the isolated / FP64 legs")
    ap.add_argument("--out", default=None, help="stream: also write the summary to this file")
    a = ap.parse_args()
    if a.stream:
        return stream_main(a)
    import torch
    torch.cuda.init()   # the runtime is initialised by torch before the server's threads use it
    t0 = time.perf_counter()
    reqs = make_requests(a.requests, a.seed, a.nmin, a.nmax, a.rate)
    print(f"# {len(reqs)} requests generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    extra = {"orthogonalization": a.orth} if a.orth else {}
    opts = server_opts(executors=a.executors, max_batch=a.max_batch, small_solver=a.small_solver,
                       wg_threads=a.wg_threads, **extra)
    sols, infos, wall, batch, wg = run_server(reqs, opts, warmup=a.warmup)
    lat = np.array([i.t_end - i.t_submit for i in infos])
    res = np.array([i.true_rel_residual for i in infos])
    out = {"requests": len(reqs), "rate": a.rate, "nmin": a.nmin, "nmax": a.nmax,
           "server": {"solves_per_s": len(reqs) / wall, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                      "p99_ms": 1e3 * float(np.percentile(lat, 99)), "max_true_rel_residual": float(res.max()),
                      "converged": int(sum(i.status == 0 for i in infos)), "batch_size": batch,
                      "wg_threads": wg,
                      "batches": len({i.batch_id for i in infos}),
                      "mean_prep_ms": 1e3 * float(np.mean([i.prep_s for i in infos])),
                      "mean_setup_ms": 1e3 * float(np.mean([i.setup_s for i in infos])),
                      "mean_solve_ms": 1e3 * float(np.mean([i.solve_s for i in infos])),
                      "mean_queue_ms": 1e3 * float(np.mean([i.t_start - i.t_submit - i.prep_s for i in infos]))}}
    print(f"# server: {out['server']}", file=sys.stderr, flush=True)
    if a.isolated:
        iso = run_isolated(reqs, server_opts(small_solver=a.small_solver, wg_threads=a.wg_threads, **extra))
        out["isolated"] = {"p50_ms": 1e3 * float(np.percentile(iso, 50)), "p99_ms": 1e3 * float(np.percentile(iso, 99))}
        out["p99_loaded_over_isolated"] = float(np.percentile(lat, 99) / np.percentile(iso, 99))
    if not a.no_fp64:
        s64, service, lat64, end = run_sequential_fp64(reqs)
        diff = max(float(np.linalg.norm(u - v) / max(np.linalg.norm(v), 1e-300)) for u, v in zip(sols, s64))
        out["fp64_sequential"] = {"solves_per_s": len(reqs) / end, "p50_ms": 1e3 * float(np.percentile(lat64, 50)),
                                  "p99_ms": 1e3 * float(np.percentile(lat64, 99)),
                                  "mean_service_ms": 1e3 * float(service.mean())}
        out["max_rel_diff_vs_fp64"] = diff
    print(json.dumps(out))


if __name__ == "__main__":
    main()
