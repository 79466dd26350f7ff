// Spectral reordering (00_CommonFiles/Reordering/reordering.c:27-77): unknowns sorted by their
// entry of the Fiedler vector (the eigenvector of the second-smallest eigenvalue of the graph
// Laplacian of the symmetric CSR pattern). The reference obtains the vector from HSL_MC73
// (mc73_fiedler, a third-party multilevel Fortran code that is not part of the reference tree in
// a buildable form); this is an independent multilevel computation of the same vector:
//   * coarsening by heavy-edge matching (vertices visited by increasing degree, ties by index)
//     until at most 128 vertices remain;
//   * the coarsest Laplacian's eigenvectors by cyclic Jacobi rotations;
//   * at every finer level, the interpolated vector refined by power iteration on sigma*I - L
//     (sigma = the Gershgorin bound 2*max degree), kept orthogonal to the constant vector.
// Everything is deterministic. The result is an ordering heuristic: its entries, not only their
// order, depend on the number of refinement iterations.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.h"

namespace {

struct WGraph {
    int n = 0;
    std::vector<int> ptr, adj;
    std::vector<double> w;   // edge weights (symmetric)
};

WGraph from_pattern(int n, const int *IA, const int *JA)
{
    std::vector<std::vector<int>> nb(n);
    for (int i = 0; i < n; ++i)
        for (int p = IA[i]; p < IA[i + 1]; ++p) {
            const int j = JA[p];
            if (j == i || j < 0 || j >= n) continue;
            nb[i].push_back(j);
            nb[j].push_back(i);
        }
    WGraph G;
    G.n = n;
    G.ptr.assign(static_cast<size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        std::sort(nb[i].begin(), nb[i].end());
        nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
        G.ptr[static_cast<size_t>(i) + 1] = G.ptr[i] + static_cast<int>(nb[i].size());
    }
    G.adj.reserve(G.ptr[n]);
    for (int i = 0; i < n; ++i) G.adj.insert(G.adj.end(), nb[i].begin(), nb[i].end());
    G.w.assign(G.adj.size(), 1.0);
    return G;
}

// Heavy-edge matching; returns the coarse graph and fine -> coarse map.
WGraph coarsen(const WGraph &G, std::vector<int> &map)
{
    const int n = G.n;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return G.ptr[a + 1] - G.ptr[a] < G.ptr[b + 1] - G.ptr[b]; });
    std::vector<int> mate(n, -1);
    for (int v : order) {
        if (mate[v] >= 0) continue;
        int best = -1;
        double bw = -1.0;
        for (int p = G.ptr[v]; p < G.ptr[v + 1]; ++p) {
            const int u = G.adj[p];
            if (mate[u] < 0 && u != v && G.w[p] > bw) {
                bw = G.w[p];
                best = u;
            }
        }
        mate[v] = best >= 0 ? best : v;
        if (best >= 0) mate[best] = v;
    }
    map.assign(n, -1);
    int nc = 0;
    for (int v = 0; v < n; ++v)
        if (map[v] < 0) {
            map[v] = nc;
            map[mate[v]] = nc;
            ++nc;
        }
    // coarse edges: summed weights, self loops dropped
    std::vector<std::vector<std::pair<int, double>>> nb(nc);
    for (int v = 0; v < n; ++v)
        for (int p = G.ptr[v]; p < G.ptr[v + 1]; ++p) {
            const int a = map[v], b = map[G.adj[p]];
            if (a != b) nb[a].emplace_back(b, G.w[p]);
        }
    WGraph C;
    C.n = nc;
    C.ptr.assign(static_cast<size_t>(nc) + 1, 0);
    for (int a = 0; a < nc; ++a) {
        auto &l = nb[a];
        std::sort(l.begin(), l.end());
        size_t k = 0;
        for (size_t q = 0; q < l.size(); ++q) {
            if (k > 0 && l[k - 1].first == l[q].first) l[k - 1].second += l[q].second;
            else l[k++] = l[q];
        }
        l.resize(k);
        C.ptr[static_cast<size_t>(a) + 1] = C.ptr[a] + static_cast<int>(k);
    }
    for (int a = 0; a < nc; ++a)
        for (auto &e : nb[a]) {
            C.adj.push_back(e.first);
            C.w.push_back(e.second);
        }
    return C;
}

void remove_mean_normalise(std::vector<double> &x)
{
    const double m = std::accumulate(x.begin(), x.end(), 0.0) / static_cast<double>(x.size());
    double s = 0.0;
    for (double &v : x) {
        v -= m;
        s += v * v;
    }
    s = std::sqrt(s);
    if (s > 0)
        for (double &v : x) v /= s;
}

// Fiedler vector of a small graph: dense Laplacian, cyclic Jacobi eigen-decomposition.
std::vector<double> fiedler_dense(const WGraph &G)
{
    const int n = G.n;
    std::vector<double> L(static_cast<size_t>(n) * n, 0.0), V(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) {
        V[static_cast<size_t>(i) * n + i] = 1.0;
        for (int p = G.ptr[i]; p < G.ptr[i + 1]; ++p) {
            L[static_cast<size_t>(i) * n + G.adj[p]] -= G.w[p];
            L[static_cast<size_t>(i) * n + i] += G.w[p];
        }
    }
    auto at = [&](int r, int c) -> double & { return L[static_cast<size_t>(r) * n + c]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += at(p, q) * at(p, q);
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (std::fabs(at(p, q)) < 1e-300) continue;
                const double theta = (at(q, q) - at(p, p)) / (2.0 * at(p, q));
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double a = at(k, p), b = at(k, q);
                    at(k, p) = c * a - s * b;
                    at(k, q) = s * a + c * b;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double a = at(p, k), b = at(q, k);
                    at(p, k) = c * a - s * b;
                    at(q, k) = s * a + c * b;
                }
                for (int k = 0; k < n; ++k) {
                    double &a = V[static_cast<size_t>(k) * n + p], &b = V[static_cast<size_t>(k) * n + q];
                    const double va = a, vb = b;
                    a = c * va - s * vb;
                    b = s * va + c * vb;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return at(a, a) < at(b, b); });
    const int col = n > 1 ? idx[1] : idx[0];
    std::vector<double> x(n);
    for (int k = 0; k < n; ++k) x[k] = V[static_cast<size_t>(k) * n + col];
    remove_mean_normalise(x);
    return x;
}

void refine(const WGraph &G, std::vector<double> &x, int iters)
{
    const int n = G.n;
    double sigma = 0.0;
    std::vector<double> deg(n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int p = G.ptr[i]; p < G.ptr[i + 1]; ++p) deg[i] += G.w[p];
        sigma = std::max(sigma, 2.0 * deg[i]);
    }
    std::vector<double> y(n);
    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < n; ++i) {
            double lx = deg[i] * x[i];
            for (int p = G.ptr[i]; p < G.ptr[i + 1]; ++p) lx -= G.w[p] * x[G.adj[p]];
            y[i] = sigma * x[i] - lx;
        }
        x.swap(y);
        remove_mean_normalise(x);
    }
}

}  // namespace

// Fiedler vector of the pattern's graph (disconnected graphs: of the whole graph's Laplacian,
// which the power iteration reaches through the components' coupling-free spectrum).
extern "C" int femgpu_fiedler(int n, const int *IA, const int *JA, int refine_iters, double *fvec)
{
    FG_CHECK(n >= 0 && (n == 0 || (IA && JA && fvec)) && refine_iters >= 0, FEMGPU_ERR_ARG,
             "femgpu_fiedler: bad arguments");
    if (n == 0) return FEMGPU_OK;
    std::vector<WGraph> levels;
    std::vector<std::vector<int>> maps;
    levels.push_back(from_pattern(n, IA, JA));
    while (levels.back().n > 128) {
        std::vector<int> m;
        WGraph C = coarsen(levels.back(), m);
        if (C.n >= levels.back().n) break;   // no matching left (e.g. no edges)
        maps.push_back(std::move(m));
        levels.push_back(std::move(C));
    }
    std::vector<double> x;
    if (levels.back().n <= 1024) {
        x = fiedler_dense(levels.back());
    } else {   // no edges left to contract: start from the index ramp
        x.resize(levels.back().n);
        for (int v = 0; v < levels.back().n; ++v) x[v] = v;
        remove_mean_normalise(x);
    }
    for (int l = static_cast<int>(levels.size()) - 2; l >= 0; --l) {
        const std::vector<int> &m = maps[l];
        std::vector<double> xf(levels[l].n);
        for (int v = 0; v < levels[l].n; ++v) xf[v] = x[m[v]];
        x.swap(xf);
        remove_mean_normalise(x);
        refine(levels[l], x, refine_iters);
    }
    // sign: the first unknown takes the non-positive side
    if (x[0] > 0)
        for (double &v : x) v = -v;
    std::copy(x.begin(), x.end(), fvec);
    return FEMGPU_OK;
}

// REORDERING_SPECTRAL (reordering.c:27-77): perm[new] = old, ascending Fiedler entries (ties by
// index).
extern "C" int femgpu_spectral(int n, const int *IA, const int *JA, int *perm)
{
    FG_CHECK(n >= 0 && (n == 0 || (IA && JA && perm)), FEMGPU_ERR_ARG, "femgpu_spectral: bad arguments");
    if (n == 0) return FEMGPU_OK;
    std::vector<double> f(n);
    FG_TRY(femgpu_fiedler(n, IA, JA, 60, f.data()));
    std::iota(perm, perm + n, 0);
    std::stable_sort(perm, perm + n, [&](int a, int b) { return f[a] < f[b]; });
    return FEMGPU_OK;
}
