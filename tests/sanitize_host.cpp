// Host-side entry points of libfemgpu (structure, mesh generation and files, Delaunay, spectral
// ordering) run under AddressSanitizer and UndefinedBehaviorSanitizer. Built by
// tests/test_host_sanitizers.py with g++ from the library's host sources; every check below is a
// structural invariant (permutations, sorted rows, slot ranges, colouring and level constraints,
// file round trips), so the sanitizers see every code path the checks drive.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "femgpu.h"

static int g_fail = 0;

#define EXPECT(cond, ...)                                                              \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "check failed at line %d: %s: ", __LINE__, #cond);    \
            std::fprintf(stderr, __VA_ARGS__);                                         \
            std::fprintf(stderr, " (%s)\n", femgpu_last_error());                      \
            ++g_fail;                                                                  \
        }                                                                              \
    } while (0)

static bool is_permutation(const std::vector<int> &p, int n)
{
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i) {
        if (p[i] < 0 || p[i] >= n || seen[p[i]]) return false;
        seen[p[i]] = 1;
    }
    return true;
}

static void check_csr(const std::vector<int> &IA, const std::vector<int> &JA, int n, int nnz, const char *what)
{
    EXPECT(IA[0] == 0 && IA[n] == nnz, "%s: IA ends %d %d nnz %d", what, IA[0], IA[n], nnz);
    for (int i = 0; i < n; ++i) {
        EXPECT(IA[i] <= IA[i + 1], "%s: IA not monotone at %d", what, i);
        for (int p = IA[i]; p < IA[i + 1]; ++p) {
            EXPECT(JA[p] >= 0 && JA[p] < n, "%s: column %d out of range in row %d", what, JA[p], i);
            if (p > IA[i]) EXPECT(JA[p - 1] < JA[p], "%s: row %d not strictly sorted", what, i);
        }
    }
}

// One problem of the structured generator through numbering, lm, CSR structure (every reordering,
// with and without the lm permutation), the EBE permutation, colouring and ILU levels.
static void structure_case(const char *problem, int mx, int my, bool right)
{
    const int nn = (mx + 1) * (my + 1), nel = 2 * mx * my;
    std::vector<double> x(nn), y(nn);
    std::vector<int> ien(3 * nel);
    EXPECT(femgpu_structured_mesh(mx, my, 0.0, 0.0, 1.0, 1.0, right ? 1 : 0, x.data(), y.data(), ien.data()) == 0,
           "%s mesh", problem);
    for (int e = 0; e < nel; ++e) {
        const int a = ien[3 * e], b = ien[3 * e + 1], c = ien[3 * e + 2];
        EXPECT(a >= 0 && a < nn && b >= 0 && b < nn && c >= 0 && c < nn, "element %d", e);
        if (a < 0 || a >= nn || b < 0 || b >= nn || c < 0 || c >= nn) return;
        const double area2 = (x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a]);
        EXPECT(std::fabs(std::fabs(area2) - 1.0 / (mx * my)) < 1e-12, "element %d area", e);
    }
    std::vector<int> mark(nn);
    const int ndof = femgpu_structured_marks(problem, nn, x.data(), y.data(), mark.data(), nullptr);
    EXPECT(ndof >= 1 && ndof <= 4, "%s ndof %d", problem, ndof);
    if (ndof < 1 || ndof > 4) return;
    std::vector<int> type(static_cast<size_t>(nn) * ndof), invP(type.size());
    EXPECT(femgpu_structured_marks(problem, nn, x.data(), y.data(), mark.data(), type.data()) == ndof, "%s marks",
           problem);
    const int neq = femgpu_number_unknowns(nn, ndof, type.data(), invP.data());
    EXPECT(neq > 0 && neq <= nn * ndof, "%s neq %d", problem, neq);
    if (neq <= 0) return;
    const int s = 3 * ndof;
    std::vector<int> lm(static_cast<size_t>(nel) * s);
    EXPECT(femgpu_fill_lm(nel, ndof, neq, ien.data(), type.data(), invP.data(), lm.data()) == 0, "fill_lm");
    for (int v : lm) EXPECT(v >= 0 && v <= neq, "lm entry %d", v);

    for (int reorder = 0; reorder <= 2; ++reorder)
        for (int permute = 0; permute <= 1; ++permute) {
            std::vector<int> lmc = lm, invc = invP;
            const int nnz = femgpu_csr_nnz(nel, s, neq, lmc.data());
            EXPECT(nnz >= neq, "nnz %d", nnz);
            if (nnz < neq) continue;
            std::vector<int> IA(neq + 1), JA(nnz), scheme(static_cast<size_t>(nel) * s * s), perm(neq);
            const int r = femgpu_csr_init(nel, s, neq, lmc.data(), reorder, permute, nn * ndof, type.data(),
                                          invc.data(), IA.data(), JA.data(), scheme.data(), perm.data());
            EXPECT(r == nnz, "csr_init %d vs %d (reorder %d)", r, nnz, reorder);
            if (r != nnz) continue;
            check_csr(IA, JA, neq, nnz, "csr_init");
            EXPECT(is_permutation(perm, neq), "csr_init perm (reorder %d)", reorder);
            for (int v : scheme) EXPECT(v >= 0 && v <= nnz, "slot %d", v);
            for (int v : lmc) EXPECT(v >= 0 && v <= neq, "permuted lm entry %d", v);
            // every element entry's slot holds the pair (row, column) of its lm entries
            if (permute)
                for (int e = 0; e < nel; ++e)
                    for (int a = 0; a < s; ++a)
                        for (int b = 0; b < s; ++b) {
                            const int i = lmc[static_cast<size_t>(e) * s + a], j = lmc[static_cast<size_t>(e) * s + b];
                            const int slot = scheme[(static_cast<size_t>(e) * s + a) * s + b];
                            if (i == neq || j == neq) {
                                EXPECT(slot == nnz, "dummy slot");
                            } else {
                                EXPECT(slot >= IA[i] && slot < IA[i + 1] && JA[slot] == j, "slot of (%d,%d)", i, j);
                            }
                        }
            if (reorder == 0 && !permute) {
                std::vector<int> p2(neq), p3(neq), lev(neq);
                EXPECT(femgpu_symrcm(neq, IA.data(), JA.data(), p2.data()) == 0, "symrcm");
                EXPECT(is_permutation(p2, neq), "symrcm perm");
                EXPECT(femgpu_spectral(neq, IA.data(), JA.data(), p3.data()) == 0, "spectral");
                EXPECT(is_permutation(p3, neq), "spectral perm");
                for (int lower = 0; lower <= 1; ++lower) {
                    const int depth = femgpu_ilu_levels(neq, IA.data(), JA.data(), lower, lev.data());
                    EXPECT(depth >= 1, "ilu depth %d", depth);
                    for (int i = 0; i < neq; ++i)
                        for (int p = IA[i]; p < IA[i + 1]; ++p) {
                            const int j = JA[p];
                            if (lower ? j < i : j > i) EXPECT(lev[i] > lev[j], "level of row %d vs %d", i, j);
                        }
                }
            }
        }

    if (ndof == 1)
        for (int reorder = 0; reorder <= 2; ++reorder) {
            std::vector<int> Perm(neq + 1), invPerm(neq + 1);
            EXPECT(femgpu_ebe_permutation(nel, neq, ien.data(), type.data(), invP.data(), reorder, Perm.data(),
                                          invPerm.data()) == 0,
                   "ebe_permutation");
            EXPECT(is_permutation(Perm, neq + 1) && Perm[neq] == neq, "Perm");
            for (int i = 0; i <= neq; ++i) EXPECT(invPerm[Perm[i]] == i, "invPerm at %d", i);
        }

    std::vector<int> colour(nel);
    const int nc = femgpu_ebe_colour(nel, s, neq, lm.data(), colour.data());
    EXPECT(nc >= 1, "colours %d", nc);
    std::vector<int> last(neq, -1), last_e(neq, -1);
    for (int c = 0; c < nc; ++c)
        for (int e = 0; e < nel; ++e) {
            if (colour[e] != c) continue;
            for (int a = 0; a < s; ++a) {
                const int i = lm[static_cast<size_t>(e) * s + a];
                if (i == neq) continue;
                EXPECT(!(last[i] == c && last_e[i] != e), "elements %d and %d share unknown %d in colour %d",
                       last_e[i], e, i, c);
                last[i] = c;
                last_e[i] = e;
            }
        }

    // mesh file round trip
    const std::string path = std::string("sanitize_") + problem + ".dat";
    EXPECT(femgpu_mesh_file_write(path.c_str(), nn, nel, ndof, x.data(), y.data(), type.data(), ien.data()) == 0,
           "write");
    femgpu_mesh_file_info info{};
    EXPECT(femgpu_mesh_file_read(path.c_str(), 0, ndof, &info, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr) == 0,
           "read sizes");
    EXPECT(info.nnodes == nn && info.nel == nel, "sizes %d %d", info.nnodes, info.nel);
    if (info.nnodes == nn && info.nel == nel) {
        std::vector<double> x2(nn), y2(nn);
        std::vector<int> t2(type.size()), i2(type.size()), ien2(ien.size()), et(nel);
        EXPECT(femgpu_mesh_file_read(path.c_str(), 0, ndof, &info, x2.data(), y2.data(), t2.data(), i2.data(),
                                     ien2.data(), et.data()) == 0,
               "read");
        EXPECT(t2 == type && ien2 == ien, "round trip of types and connectivity");
        double dx = 0;
        for (int i = 0; i < nn; ++i) dx = std::max(dx, std::fabs(x2[i] - x[i]) + std::fabs(y2[i] - y[i]));
        EXPECT(dx < 1e-6, "coordinates %g", dx);
    }
    std::remove(path.c_str());
}

static void delaunay_case(int n, int dups, unsigned seed, bool grid)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> x, y;
    if (grid) {
        const int m = static_cast<int>(std::sqrt(static_cast<double>(n)));
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                x.push_back(i / double(m - 1));
                y.push_back(j / double(m - 1));
            }
    } else {
        for (int i = 0; i < n; ++i) {
            x.push_back(u(gen));
            y.push_back(u(gen));
        }
    }
    const int base = static_cast<int>(x.size());
    for (int k = 0; k < dups; ++k) {
        const int i = static_cast<int>(gen() % base);
        x.push_back(x[i]);
        y.push_back(y[i]);
    }
    const int N = static_cast<int>(x.size()), cap = 2 * N;
    std::vector<int> tri(3 * static_cast<size_t>(cap));
    int nt = 0, nd = -1;
    EXPECT(femgpu_delaunay(N, x.data(), y.data(), cap, &nt, tri.data(), &nd) == 0, "delaunay n %d", N);
    EXPECT(nd == dups, "duplicates %d vs %d", nd, dups);
    EXPECT(nt >= 1 && nt <= 2 * base - 5, "triangles %d for %d points", nt, base);
    double area = 0;
    for (int t = 0; t < nt; ++t) {
        const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        EXPECT(a >= 0 && a < N && b >= 0 && b < N && c >= 0 && c < N, "triangle %d", t);
        if (a < 0 || a >= N || b < 0 || b >= N || c < 0 || c >= N) return;
        const double a2 = (x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a]);
        EXPECT(a2 > 0, "triangle %d not counter-clockwise", t);
        area += 0.5 * a2;
    }
    if (grid) EXPECT(std::fabs(area - 1.0) < 1e-12, "grid area %g", area);
}

int main()
{
    structure_case("CONE", 9, 7, true);
    structure_case("TESTE", 12, 12, false);
    structure_case("CAVITY", 8, 8, true);
    structure_case("OBLIQUO", 10, 5, true);
    delaunay_case(3000, 17, 1u, false);
    delaunay_case(400, 0, 2u, true);
    delaunay_case(10, 3, 3u, false);
    // invalid arguments come back as status codes
    EXPECT(femgpu_symrcm(-1, nullptr, nullptr, nullptr) != 0, "symrcm rejects n < 0");
    EXPECT(femgpu_delaunay(2, nullptr, nullptr, 0, nullptr, nullptr, nullptr) != 0, "delaunay rejects nulls");
    std::printf("%s: %d failed checks\n", g_fail ? "FAIL" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
