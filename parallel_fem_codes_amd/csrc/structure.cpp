// Host-side integer work of the hot path: unknown numbering, element maps, the CSR structure with
// its per-element slot table and reordering, the EBE permutation, element colouring, ILU(0)
// dependency levels and the chunked sweep schedules. Everything here must be bit-exact with the
// reference's integer results (see the citations at each entry point in include/femgpu.h).
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <queue>
#include <vector>

#include "common.h"

namespace {

// Sorted, de-duplicated adjacency of the rows [0, nrows) induced by element maps: every entry
// lm[e*s+a] in range couples with every entry lm[e*s+b] in range of the same element.
// Row lists come out ascending, which is the order csr_List_insertA keeps its lists in.
void element_pattern(int nel, int s, int nrows, const int *lm, std::vector<int> &IA,
                     std::vector<int> &JA)
{
    std::vector<int> cnt(static_cast<size_t>(nrows) + 1, 0);
    for (int e = 0; e < nel; ++e) {
        const int *L = lm + static_cast<size_t>(e) * s;
        int inrange = 0;
        for (int b = 0; b < s; ++b) inrange += (L[b] >= 0 && L[b] < nrows);
        for (int a = 0; a < s; ++a)
            if (L[a] >= 0 && L[a] < nrows) cnt[L[a] + 1] += inrange;
    }
    for (int i = 0; i < nrows; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> buf(cnt[nrows]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < nel; ++e) {
        const int *L = lm + static_cast<size_t>(e) * s;
        for (int a = 0; a < s; ++a) {
            if (L[a] < 0 || L[a] >= nrows) continue;
            for (int b = 0; b < s; ++b)
                if (L[b] >= 0 && L[b] < nrows) buf[fill[L[a]]++] = L[b];
        }
    }
    IA.assign(static_cast<size_t>(nrows) + 1, 0);
    JA.clear();
    JA.reserve(buf.size() / 2 + 1);
    // per row: drop repeated columns with a last-row marker, then sort the (few) distinct ones
    std::vector<int> mark(static_cast<size_t>(nrows), -1);
    for (int i = 0; i < nrows; ++i) {
        const size_t start = JA.size();
        for (int q = cnt[i]; q < cnt[i + 1]; ++q)
            if (mark[buf[q]] != i) {
                mark[buf[q]] = i;
                JA.push_back(buf[q]);
            }
        std::sort(JA.begin() + static_cast<std::ptrdiff_t>(start), JA.end());
        IA[i + 1] = static_cast<int>(JA.size());
    }
}

// ---- Reverse Cuthill-McKee with the exact tie-breaking of symrcm.c ----------------------------
// BFS records carry (id, degree, distance). The neighbour heap is a binary min-heap on degree whose
// insert sifts up on strict '<' and whose removal moves the last record to the root and sifts down
// preferring the left child, then the right one, on strict '<' (symrcm.c:505-567). These rules fix
// the order of equal-degree neighbours and therefore the permutation.
struct Rec {
    int id, deg, dist;
};

struct DegreeHeap {
    std::vector<Rec> h;
    int size = 0;
    explicit DegreeHeap(int cap) : h(static_cast<size_t>(cap) + 1) {}
    void push(const Rec &r)
    {
        int i = size++;
        h[i] = r;
        while (i > 0) {
            int p = (i - 1) / 2;
            if (!(h[i].deg < h[p].deg)) break;
            std::swap(h[i], h[p]);
            i = p;
        }
    }
    Rec pop()
    {
        Rec top = h[0];
        h[0] = h[--size];
        for (int j = 0;;) {
            int l = 2 * j + 1, r = l + 1, m = j;
            if (l < size && h[l].deg < h[m].deg) m = l;
            if (r < size && h[r].deg < h[m].deg) m = r;
            if (m == j) break;
            std::swap(h[j], h[m]);
            j = m;
        }
        return top;
    }
};

// Pattern plus its transpose; neighbours of i are the union of row i of both, ascending, an index
// present in both visited once (symrcm.c:116-176, :386-459).
struct SymGraph {
    const int *IA, *JA;
    std::vector<int> TP, TJ;
    int n;
    SymGraph(int n_, const int *IA_, const int *JA_) : IA(IA_), JA(JA_), n(n_)
    {
        const int nnz = IA[n];
        TP.assign(static_cast<size_t>(n) + 1, 0);
        TJ.assign(static_cast<size_t>(nnz) + 1, 0);
        for (int k = 0; k < nnz; ++k) TP[JA[k] + 1]++;
        for (int i = 0; i < n; ++i) TP[i + 1] += TP[i];
        std::vector<int> pos(TP.begin(), TP.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int k = IA[r]; k < IA[r + 1]; ++k) TJ[pos[JA[k]]++] = r;
    }
    template <typename F>
    void for_each_neighbour(int i, F &&f) const
    {
        int a = IA[i], ae = IA[i + 1], b = TP[i], be = TP[i + 1];
        while (a < ae || b < be) {
            if (a == ae) { f(TJ[b++]); continue; }
            if (b == be) { f(JA[a++]); continue; }
            int u = JA[a], v = TJ[b];
            if (u <= v) {
                f(u);
                ++a;
                if (u == v) ++b;
            } else {
                f(v);
                ++b;
            }
        }
    }
};

// Degrees as calc_degrees counts them (symrcm.c:305-350): every stored entry counts once for its
// row index, and an entry without a mirrored partner also counts for its column.
int rcm_degrees(int n, const int *IA, const int *JA, std::vector<int> &D)
{
    D.assign(n, 0);
    int maxd = 0;
    for (int j = 0; j < n; ++j)
        for (int p = IA[j]; p < IA[j + 1]; ++p) {
            const int k = JA[p];
            maxd = std::max(maxd, ++D[k]);
            if (k == j) continue;
            bool mirrored = false;
            for (int q = IA[k]; q < IA[k + 1]; ++q) {
                if (JA[q] == j) { mirrored = true; break; }
                if (JA[q] > j) break;
            }
            if (!mirrored) maxd = std::max(maxd, ++D[j]);
        }
    return maxd;
}

// Pseudo-peripheral start (find_starting_node, symrcm.c:352-483): repeated BFS; the candidate is
// replaced by any dequeued record that is farther, or (not itself and) of larger degree.
int rcm_start_node(const SymGraph &G, const std::vector<int> &D, int start)
{
    const int n = G.n;
    std::vector<Rec> Q(static_cast<size_t>(n) + 1);
    std::vector<char> seen(n, 0);
    int qh = 0, qt = 0;
    Rec x{start, D[start], 0};
    auto enq = [&](const Rec &r) { Q[qt] = r; qt = (qt + 1) % (n + 1); };
    enq(x);
    seen[start] = 1;
    int max_dist = 0;
    for (;;) {
        while (qh != qt) {
            Rec v = Q[qh];
            qh = (qh + 1) % (n + 1);
            if (v.dist > x.dist || (v.id != x.id && v.deg > x.deg)) x = v;
            G.for_each_neighbour(v.id, [&](int r) {
                if (!seen[r]) {
                    enq(Rec{r, D[r], v.dist + 1});
                    seen[r] = 1;
                }
            });
        }
        if (max_dist < x.dist) {
            max_dist = x.dist;
            std::fill(seen.begin(), seen.end(), 0);
            seen[x.id] = 1;
            x.dist = 0;
            qh = qt = 0;
            enq(x);
        } else {
            break;
        }
    }
    return x.id;
}

}  // namespace

extern "C" int femgpu_symrcm(int n, const int *IA, const int *JA, int *perm)
{
    FG_CHECK(n >= 0 && (n == 0 || (IA && JA && perm)), FEMGPU_ERR_ARG, "femgpu_symrcm: bad arguments");
    if (n == 0) return FEMGPU_OK;
    std::vector<int> D;
    const int maxd = rcm_degrees(n, IA, JA, D);
    if (maxd == 0) {
        std::iota(perm, perm + n, 0);
        return FEMGPU_OK;
    }
    SymGraph G(n, IA, JA);
    DegreeHeap heap(2 * maxd + 2);
    std::vector<Rec> Q(static_cast<size_t>(n) + 1);
    std::vector<char> visited(n, 0);
    int qh = 0, qt = 0, c = -1, scan = 0;
    do {
        while (scan < n && visited[scan]) ++scan;   // first unvisited index (symrcm.c:75-77)
        Rec v{rcm_start_node(G, D, scan), 0, 0};
        v.deg = D[v.id];
        visited[v.id] = 1;
        Q[qt] = v;
        qt = (qt + 1) % (n + 1);
        while (qh != qt) {
            v = Q[qh];
            qh = (qh + 1) % (n + 1);
            perm[++c] = v.id;
            G.for_each_neighbour(v.id, [&](int r) {
                if (!visited[r]) {
                    heap.push(Rec{r, D[r], v.dist + 1});
                    visited[r] = 1;
                }
            });
            while (heap.size) {
                Q[qt] = heap.pop();
                qt = (qt + 1) % (n + 1);
            }
        }
        scan = 0;
    } while (c + 1 < n);
    std::reverse(perm, perm + n);   // symrcm.c:231-235
    return FEMGPU_OK;
}

extern "C" int femgpu_number_unknowns(int nnodes, int ndof, const int *type, int *invP_id)
{
    FG_CHECK(nnodes >= 0 && ndof > 0 && type && invP_id, FEMGPU_ERR_ARG,
             "femgpu_number_unknowns: bad arguments");
    int neq = 0;
    const size_t total = static_cast<size_t>(nnodes) * ndof;
    for (size_t k = 0; k < total; ++k) invP_id[k] = (type[k] == 1) ? neq++ : -1;
    return neq;
}

extern "C" int femgpu_fill_lm(int nel, int ndof, int NEQ, const int *ien, const int *type,
                              const int *invP_id, int *lm)
{
    FG_CHECK(nel >= 0 && ndof > 0 && ien && type && invP_id && lm, FEMGPU_ERR_ARG,
             "femgpu_fill_lm: bad arguments");
    const int s = 3 * ndof;
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 3; ++a) {
            const size_t base = static_cast<size_t>(ien[3 * e + a]) * ndof;
            for (int d = 0; d < ndof; ++d)
                lm[static_cast<size_t>(e) * s + a * ndof + d] = type[base + d] ? invP_id[base + d] : NEQ;
        }
    return FEMGPU_OK;
}

extern "C" int femgpu_csr_nnz(int nel, int s, int neq, const int *lm)
{
    FG_CHECK(nel >= 0 && s > 0 && neq >= 0 && lm, FEMGPU_ERR_ARG, "femgpu_csr_nnz: bad arguments");
    std::vector<int> IA, JA;
    element_pattern(nel, s, neq, lm, IA, JA);
    return static_cast<int>(JA.size());
}

// csr_Initialization (ST/csr_Initialization.c:5-571) into vectors: the sorted element pattern, the
// slot of every element entry, the reordering and the permuted structure (see femgpu_csr_init).
int fg::csr_init_vec(int nel, int s, int neq, int *lm, int reorder, int permute_lm, int ndofs_total, const int *type,
                     int *invP_id, std::vector<int> &IA, std::vector<int> &JA, std::vector<int> &scheme,
                     std::vector<int> &perm)
{
    FG_CHECK(nel >= 0 && s > 0 && neq >= 0 && lm, FEMGPU_ERR_ARG, "femgpu_csr_init: bad arguments");
    FG_CHECK(reorder >= 0 && reorder <= 2, FEMGPU_ERR_ARG,
             "femgpu_csr_init: reorder must be 0 (NOT), 1 (SYMRCM) or 2 (Spectral)");
    std::vector<int> ia, ja;
    element_pattern(nel, s, neq, lm, ia, ja);
    const int nnz = static_cast<int>(ja.size());
    // Slot of every element entry in the unpermuted structure (csr_Initialization.c:101-191).
    const size_t ss = static_cast<size_t>(s) * s;
    scheme.resize(static_cast<size_t>(nel) * ss);
    for (int e = 0; e < nel; ++e) {
        const int *L = lm + static_cast<size_t>(e) * s;
        for (int a = 0; a < s; ++a)
            for (int b = 0; b < s; ++b) {
                int slot = nnz;
                if (L[a] >= 0 && L[a] < neq && L[b] >= 0 && L[b] < neq) {
                    auto first = ja.begin() + ia[L[a]], last = ja.begin() + ia[L[a] + 1];
                    slot = static_cast<int>(std::lower_bound(first, last, L[b]) - ja.begin());
                }
                scheme[e * ss + a * s + b] = slot;
            }
    }
    // Reordering (reordering.c:3-22), then rows and columns permuted with per-row re-sort,
    // carrying the original slot of every entry (reordering.c:101-167).
    perm.resize(neq);
    std::iota(perm.begin(), perm.end(), 0);
    if (reorder == 1) FG_TRY(femgpu_symrcm(neq, ia.data(), ja.data(), perm.data()));
    if (reorder == 2) FG_TRY(femgpu_spectral(neq, ia.data(), ja.data(), perm.data()));
    std::vector<int> inv(static_cast<size_t>(neq) + 1);
    for (int i = 0; i < neq; ++i) inv[perm[i]] = i;
    inv[neq] = neq;
    std::vector<int> newslot(static_cast<size_t>(nnz) + 1);
    newslot[nnz] = nnz;
    IA.assign(static_cast<size_t>(neq) + 1, 0);
    JA.resize(nnz);
    std::vector<std::pair<int, int>> row;   // (new column, old slot)
    int k = 0;
    for (int i = 0; i < neq; ++i) {
        const int old = perm[i];
        row.clear();
        for (int p = ia[old]; p < ia[old + 1]; ++p) row.emplace_back(inv[ja[p]], p);
        std::sort(row.begin(), row.end());
        for (auto &cs : row) {
            JA[k] = cs.first;
            newslot[cs.second] = k;
            ++k;
        }
        IA[i + 1] = k;
    }
    for (size_t q = 0; q < static_cast<size_t>(nel) * ss; ++q) scheme[q] = newslot[scheme[q]];
    if (permute_lm)
        for (size_t q = 0; q < static_cast<size_t>(nel) * s; ++q)
            if (lm[q] >= 0 && lm[q] < neq) lm[q] = inv[lm[q]];
    if (invP_id && type)
        for (int q = 0; q < ndofs_total; ++q)
            if (type[q] == 1) invP_id[q] = inv[invP_id[q]];
    return nnz;
}

extern "C" int femgpu_csr_init(int nel, int s, int neq, int *lm, int reorder, int permute_lm,
                               int ndofs_total, const int *type, int *invP_id, int *IA, int *JA,
                               int *scheme, int *perm)
{
    FG_CHECK(nel >= 0 && s > 0 && neq >= 0 && lm && IA && JA && scheme && perm, FEMGPU_ERR_ARG,
             "femgpu_csr_init: bad arguments");
    std::vector<int> ia, ja, sch, pm;
    const int nnz = fg::csr_init_vec(nel, s, neq, lm, reorder, permute_lm, ndofs_total, type, invP_id, ia, ja, sch, pm);
    if (nnz < 0) return nnz;
    std::copy(ia.begin(), ia.end(), IA);
    std::copy(ja.begin(), ja.end(), JA);
    std::copy(sch.begin(), sch.end(), scheme);
    std::copy(pm.begin(), pm.end(), perm);
    return nnz;
}

extern "C" int femgpu_ebe_permutation(int nel, int NEQ, const int *ien, const int *type,
                                      const int *invP_id, int reorder, int *Perm, int *invPerm)
{
    FG_CHECK(nel >= 0 && NEQ >= 0 && ien && type && invP_id && Perm && invPerm, FEMGPU_ERR_ARG,
             "femgpu_ebe_permutation: bad arguments");
    FG_CHECK(reorder >= 0 && reorder <= 2, FEMGPU_ERR_ARG, "femgpu_ebe_permutation: bad reorder");
    std::vector<int> lmu(static_cast<size_t>(nel) * 3);
    for (size_t q = 0; q < lmu.size(); ++q) lmu[q] = type[ien[q]] ? invP_id[ien[q]] : -1;
    std::iota(Perm, Perm + NEQ, 0);
    Perm[NEQ] = NEQ;
    if (reorder != 0) {
        std::vector<int> IA, JA;
        element_pattern(nel, 3, NEQ, lmu.data(), IA, JA);
        if (reorder == 1) FG_TRY(femgpu_symrcm(NEQ, IA.data(), JA.data(), Perm));
        else FG_TRY(femgpu_spectral(NEQ, IA.data(), JA.data(), Perm));
    }
    for (int i = 0; i < NEQ; ++i) invPerm[Perm[i]] = i;
    invPerm[NEQ] = NEQ;
    return FEMGPU_OK;
}

extern "C" int femgpu_ebe_colour(int nel, int s, int NEQ, const int *lm, int *colour)
{
    FG_CHECK(nel >= 0 && s > 0 && lm && colour, FEMGPU_ERR_ARG, "femgpu_ebe_colour: bad arguments");
    // Colours used so far at each unknown, as a bit set per 64 colours (extended on demand).
    int words = 1;
    std::vector<uint64_t> used(static_cast<size_t>(NEQ) + 1, 0);
    int ncol = 0;
    for (int e = 0; e < nel; ++e) {
        const int *L = lm + static_cast<size_t>(e) * s;
        int c = 0;
        for (;; ++c) {
            if (c / 64 >= words) {
                // widen the bit sets
                std::vector<uint64_t> wider(static_cast<size_t>(NEQ + 1) * (words + 1), 0);
                for (int u = 0; u <= NEQ; ++u)
                    for (int w = 0; w < words; ++w) wider[static_cast<size_t>(u) * (words + 1) + w] = used[static_cast<size_t>(u) * words + w];
                used.swap(wider);
                ++words;
            }
            bool clash = false;
            for (int a = 0; a < s && !clash; ++a)
                if (L[a] >= 0 && L[a] < NEQ)
                    clash = (used[static_cast<size_t>(L[a]) * words + c / 64] >> (c % 64)) & 1u;
            if (!clash) break;
        }
        colour[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < s; ++a)
            if (L[a] >= 0 && L[a] < NEQ) used[static_cast<size_t>(L[a]) * words + c / 64] |= (uint64_t(1) << (c % 64));
    }
    return ncol;
}

extern "C" int femgpu_ilu_levels(int n, const int *IA, const int *JA, int lower, int *level)
{
    FG_CHECK(n >= 0 && IA && JA && level, FEMGPU_ERR_ARG, "femgpu_ilu_levels: bad arguments");
    int depth = 0;
    if (lower) {
        for (int i = 0; i < n; ++i) {
            int lv = 0;
            for (int p = IA[i]; p < IA[i + 1] && JA[p] < i; ++p) lv = std::max(lv, level[JA[p]] + 1);
            level[i] = lv;
            depth = std::max(depth, lv + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int lv = 0;
            for (int p = IA[i + 1] - 1; p >= IA[i] && JA[p] > i; --p) lv = std::max(lv, level[JA[p]] + 1);
            level[i] = lv;
            depth = std::max(depth, lv + 1);
        }
    }
    return depth;
}

namespace fg {

}  // namespace fg

namespace fg {

// Packed sweep plan. Sweep index s runs over rows in elimination order (row s for the lower
// factor, row n-1-s for the upper one), so both directions are "lower" in s. Rows are scheduled by
// (level, s) and cut into chunks along level bands: a chunk holds whole consecutive levels (a level
// wider than a chunk is split), so only the first level of a chunk reads rows of earlier chunks.
int build_packed_plan(femgpu_ctx *ctx, int n, const int *IA, const int *JA, const int *diag, bool upper,
                      PackedSweep &P)
{
    auto row_at = [&](int s) { return upper ? n - 1 - s : s; };
    auto sidx = [&](int row) { return upper ? n - 1 - row : row; };
    auto ent_begin = [&](int row) { return upper ? diag[row] + 1 : IA[row]; };
    auto ent_end = [&](int row) { return upper ? IA[row + 1] : diag[row]; };
    auto nent = [&](int s) { return ent_end(row_at(s)) - ent_begin(row_at(s)); };
    std::vector<int> lev(n);
    int depth = 0;
    for (int s = 0; s < n; ++s) {
        const int r = row_at(s);
        int l = 0;
        for (int p = ent_begin(r); p < ent_end(r); ++p) l = std::max(l, lev[sidx(JA[p])] + 1);
        lev[s] = l;
        depth = std::max(depth, l + 1);
    }
    P.n = n;
    P.depth = depth;
    // global schedule by (level, s): counting sort on the level
    std::vector<int> lptr(static_cast<size_t>(depth) + 1, 0), sched(n);
    for (int s = 0; s < n; ++s) lptr[lev[s] + 1]++;
    for (int l = 0; l < depth; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int s = 0; s < n; ++s) sched[fill[lev[s]]++] = s;
    }
    // chunking for a rows-per-chunk bound Rb: whole levels while they fit, wide levels split
    auto make_chunks = [&](int Rb, std::vector<int> &cptr, int &emax) {
        cptr.assign(1, 0);
        emax = 0;
        int start = 0, ents = 0;
        for (int l = 0; l < depth; ++l) {
            const int a = lptr[l], b = lptr[l + 1];
            if (b - start > Rb && a > start) {   // this level does not fit: close the chunk
                cptr.push_back(a);
                emax = std::max(emax, ents);
                start = a;
                ents = 0;
            }
            for (int q = a; q < b; ++q) {
                if (q - start == Rb) {   // a level wider than a chunk
                    cptr.push_back(q);
                    emax = std::max(emax, ents);
                    start = q;
                    ents = 0;
                }
                ents += nent(sched[q]);
            }
        }
        if (n > start || n == 0) cptr.push_back(n);
        emax = std::max(emax, ents);
    };
    static const int cand[] = {6144, 4096, 3072, 2048, 1536, 1024, 768, 512, 256, 128, 64};
    const size_t limit = ctx->lds_limit > 1024 ? ctx->lds_limit - 1024 : ctx->lds_limit;
    std::vector<int> cptr;
    for (int Rc : cand) {
        int em = 0;
        make_chunks(Rc, cptr, em);
        PackedSweep probe;
        probe.R = Rc;
        probe.emax = std::max(em, 1);
        if (packed_lds_bytes<double>(probe) <= limit || Rc == 64) {
            P.R = Rc;
            P.emax = probe.emax;
            break;
        }
    }
    P.nchunks = static_cast<int>(cptr.size()) - 1;
    std::vector<int> row_of(n), pos(n), chunk_of(n), ent_ptr(static_cast<size_t>(n) + 1),
        grp_ptr(static_cast<size_t>(P.nchunks) + 1);
    std::vector<int> grp_range, grp_flags, code, src;   // grp_range: (start, end) pairs in chunk positions
    std::vector<unsigned char> exported(n, 0);
    for (int c = 0; c < P.nchunks; ++c)
        for (int q = cptr[c]; q < cptr[c + 1]; ++q) {
            row_of[q] = row_at(sched[q]);
            pos[sched[q]] = q - cptr[c];
            chunk_of[sched[q]] = c;
        }
    code.reserve(static_cast<size_t>(IA[n]) / 2 + 1);
    src.reserve(static_cast<size_t>(IA[n]) / 2 + 1);
    for (int c = 0; c < P.nchunks; ++c) {
        const int q0 = cptr[c], q1 = cptr[c + 1];
        grp_ptr[c] = static_cast<int>(grp_flags.size());
        int cur = -1;
        for (int q = q0; q < q1; ++q) {
            const int k = q - q0, r = row_of[q], l = lev[sidx(r)];
            if (l != cur) {
                if (cur >= 0) grp_range.push_back(k);
                grp_range.push_back(k);
                grp_flags.push_back(0);
                cur = l;
            }
            ent_ptr[q] = static_cast<int>(code.size());
            // the reference's per-row order: ascending columns (ILUp_precond.c:26-44)
            for (int p = ent_begin(r); p < ent_end(r); ++p) {
                const int dr = JA[p], ds = sidx(dr);
                if (chunk_of[ds] == c) {
                    code.push_back(pos[ds]);
                } else {
                    code.push_back(-(dr + 1));
                    grp_flags.back() |= 1;
                    exported[cptr[chunk_of[ds]] + pos[ds]] = 1;
                }
                src.push_back(p);
            }
        }
        grp_range.push_back(q1 - q0);
    }
    ent_ptr[n] = static_cast<int>(code.size());
    grp_ptr[P.nchunks] = static_cast<int>(grp_flags.size());
    // Self-check of the order the kernel relies on to make progress: a local dependency sits in
    // an earlier level group of the same chunk, a remote one in an earlier chunk.
    for (int c = 0; c < P.nchunks; ++c)
        for (int g = grp_ptr[c]; g < grp_ptr[c + 1]; ++g)
            for (int k = grp_range[2 * g]; k < grp_range[2 * g + 1]; ++k)
                for (int e = ent_ptr[cptr[c] + k]; e < ent_ptr[cptr[c] + k + 1]; ++e) {
                    const int cd = code[e];
                    const bool ok = cd >= 0 ? cd < grp_range[2 * g] : chunk_of[sidx(-cd - 1)] < c;
                    FG_CHECK(ok, FEMGPU_ERR_STATE, "packed sweep plan: dependency order violated (chunk %d)", c);
                }
    FG_CHECK(P.nchunks == 0 || *std::max_element(cptr.begin() + 1, cptr.end()) <= n, FEMGPU_ERR_STATE,
             "packed sweep plan: bad chunks");
    for (int c = 0; c < P.nchunks; ++c)
        FG_CHECK(cptr[c + 1] - cptr[c] <= P.R, FEMGPU_ERR_STATE, "packed sweep plan: chunk %d too large", c);
    P.ngroups = static_cast<int>(grp_flags.size());
    P.E = static_cast<int>(code.size());
    hipStream_t st = ctx->stream;
    FG_TRY(P.chunk_ptr.upload(cptr.data(), cptr.size(), st));
    FG_TRY(P.lvl_ptr.upload(lptr.data(), lptr.size(), st));
    FG_TRY(P.row_of.upload(row_of.data(), n, st));
    FG_TRY(P.ent_ptr.upload(ent_ptr.data(), ent_ptr.size(), st));
    FG_TRY(P.ent_code.upload(code.data(), code.size(), st));
    FG_TRY(P.ent_src.upload(src.data(), src.size(), st));
    FG_TRY(P.grp_ptr.upload(grp_ptr.data(), grp_ptr.size(), st));
    FG_TRY(P.grp_start.upload(grp_range.data(), grp_range.size(), st));
    FG_TRY(P.grp_flags.upload(grp_flags.data(), grp_flags.size(), st));
    FG_TRY(P.exported.upload(exported.data(), exported.size(), st));
    FG_TRY(P.val64.alloc(static_cast<size_t>(P.E) + 1));
    FG_TRY(P.val32.alloc(static_cast<size_t>(P.E) + 1));
    FG_TRY(P.dinv64.alloc(static_cast<size_t>(n) + 1));
    FG_TRY(P.dinv32.alloc(static_cast<size_t>(n) + 1));
    FG_TRY(P.ctl.alloc(4));
    FG_TRY(P.mbox.alloc(2 * static_cast<size_t>(n) + 2));
    FG_TRY(P.mbox32.alloc(static_cast<size_t>(n) + 1));
    const unsigned ctl0[4] = {0u, 0u, 1u, 0u};   // ticket, finished, epoch (FP32 mailboxes start at epoch 0)
    FG_HIP(hipMemcpyAsync(P.ctl.p, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, st));
    {
        std::vector<unsigned long long> sentinel(2 * static_cast<size_t>(n) + 2, 0xfff8f00df00df00dull);
        FG_HIP(hipMemcpyAsync(P.mbox.p, sentinel.data(), sizeof(unsigned long long) * sentinel.size(),
                              hipMemcpyHostToDevice, st));
        FG_HIP(hipStreamSynchronize(st));
    }
    FG_HIP(hipMemsetAsync(P.mbox32.p, 0, sizeof(unsigned long long) * (static_cast<size_t>(n) + 1), st));
    FG_HIP(hipStreamSynchronize(st));
    return FEMGPU_OK;
}

// Update lists of the numeric ILU(0) (ILUp_precond_setup.c:84-166 with p = 0): for row i and each
// L entry t (column jrow, ascending), every U entry k of row jrow whose column is in row i's
// pattern updates that slot (or the pivot when the column is i), in U order.
size_t stream_lds_bytes(const StreamSweep &P, size_t elem)
{
    return elem * (static_cast<size_t>(P.W) + 2 + P.L) +
           sizeof(int) * (static_cast<size_t>(P.nsteps) + 2 * kStreamPad + 1);
}

int build_stream_plan(femgpu_ctx *ctx, int n, const int *IA, const int *JA, const int *diag, bool upper,
                      StreamSweep &P)
{
    P.ok = false;
    P.n = n;
    P.K = P.W = P.L = 0;
    if (n == 0) return FEMGPU_OK;
    auto ent_begin = [&](int row) { return upper ? diag[row] + 1 : IA[row]; };
    auto ent_end = [&](int row) { return upper ? IA[row + 1] : diag[row]; };
    // dependency levels in sweep order (ascending rows for L, descending for U)
    std::vector<int> lev(n);
    int depth = 0, maxdeg = 0;
    for (int s = 0; s < n; ++s) {
        const int r = upper ? n - 1 - s : s;
        int l = 0;
        for (int p = ent_begin(r); p < ent_end(r); ++p) l = std::max(l, lev[JA[p]] + 1);
        lev[r] = l;
        depth = std::max(depth, l + 1);
        maxdeg = std::max(maxdeg, ent_end(r) - ent_begin(r));
    }
    P.depth = depth;
    int K = 0;
    for (int k : {4, 8, 16})
        if (maxdeg <= k) { K = k; break; }
    if (K == 0) return FEMGPU_OK;   // rows too long for the register-staged kernel
    // schedule: rows by (level, sweep order); levels cut into steps of at most kStreamThreads rows
    std::vector<int> lptr(static_cast<size_t>(depth) + 1, 0), row_of(n), pos(n);
    for (int r = 0; r < n; ++r) lptr[lev[r] + 1]++;
    for (int l = 0; l < depth; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int s = 0; s < n; ++s) {
            const int r = upper ? n - 1 - s : s;
            row_of[fill[lev[r]]++] = r;
        }
    }
    for (int q = 0; q < n; ++q) pos[row_of[q]] = q;
    std::vector<int> stepq(1, 0), step_beg(n), step_end(n), step_of(n);
    for (int l = 0; l < depth; ++l)
        for (int a = lptr[l]; a < lptr[l + 1]; a += kStreamThreads) {
            const int b = std::min(lptr[l + 1], a + kStreamThreads);
            const int st = static_cast<int>(stepq.size()) - 1;
            stepq.push_back(b);
            for (int q = a; q < b; ++q) {
                step_beg[q] = a;
                step_end[q] = b;
                step_of[q] = st;
            }
        }
    // kStreamPad empty steps in front (the kernel's pipeline fills on them), empty steps at the tail
    const int nsteps_real = static_cast<int>(stepq.size()) - 1;
    const int nsteps = kStreamPad + (nsteps_real + kStreamPad - 1) / kStreamPad * kStreamPad;
    stepq.insert(stepq.begin(), kStreamPad, 0);
    stepq.resize(static_cast<size_t>(nsteps) + 2 * kStreamPad + 1, n);
    // ring size: a dependency at position qj read by a row of a step ending at e must not be
    // overwritten before e, i.e. qj + W >= e
    long long wreq = 1;
    for (int q = 0; q < n; ++q) {
        const int r = row_of[q];
        for (int p = ent_begin(r); p < ent_end(r); ++p) wreq = std::max<long long>(wreq, step_end[q] - pos[JA[p]]);
    }
    int W = 64;
    while (W < wreq && W < kStreamRingMax) W <<= 1;
    // values read beyond the ring's reach: a long-lived slot from the step that writes the value to
    // the last step that reads it (slots handed out in order of the writing step, a slot free again
    // for a value written after its last reader's step)
    std::vector<int> last_read(n, -1);   // by schedule position of the value: last reading step
    for (int q = 0; q < n; ++q) {
        const int r = row_of[q];
        for (int p = ent_begin(r); p < ent_end(r); ++p) {
            const int qj = pos[JA[p]];
            if (step_end[q] - qj > W) last_read[qj] = std::max(last_read[qj], step_of[q]);
        }
    }
    std::vector<int> lslot(n, -1);
    int L = 0;
    {
        std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<>> busy;
        std::vector<int> free_slots;
        for (int q = 0; q < n; ++q) {   // positions in step order
            if (last_read[q] < 0) continue;
            while (!busy.empty() && busy.top().first < step_of[q]) {
                free_slots.push_back(busy.top().second);
                busy.pop();
            }
            int sl;
            if (!free_slots.empty()) {
                sl = free_slots.back();
                free_slots.pop_back();
            } else {
                sl = L++;
            }
            lslot[q] = sl;
            busy.push({last_read[q], sl});
        }
    }
    if (L > kStreamLongMax || W + 2 + L > 0xffff) return FEMGPU_OK;
    P.K = K;
    P.W = W;
    P.L = L;
    P.nsteps = nsteps;
    P.steps = nsteps_real;
    const size_t limit = ctx->lds_limit > 1024 ? ctx->lds_limit - 1024 : ctx->lds_limit;
    if (stream_lds_bytes(P, sizeof(double)) > limit) return FEMGPU_OK;
    // padded row data in schedule order, entries in the serial sweep's order
    std::vector<unsigned short> slot(static_cast<size_t>(n) * K, static_cast<unsigned short>(W));
    std::vector<int> src(static_cast<size_t>(n) * K, -1);
    for (int q = 0; q < n; ++q) {
        const int r = row_of[q];
        int k = 0;
        for (int p = ent_begin(r); p < ent_end(r); ++p, ++k) {
            const int qj = pos[JA[p]];
            if (qj >= step_beg[q]) {
                set_error("stream plan: dependency of row %d is not scheduled before it", r);
                return FEMGPU_ERR_ARG;
            }
            const bool far = step_end[q] - qj > W;
            slot[static_cast<size_t>(q) * K + k] =
                static_cast<unsigned short>(far ? W + 2 + lslot[qj] : (qj & (W - 1)));
            src[static_cast<size_t>(q) * K + k] = p;
        }
    }
    hipStream_t st = ctx->stream;
    FG_TRY(P.stepq.upload(stepq.data(), stepq.size(), st));
    FG_TRY(P.row_of.upload(row_of.data(), n, st));
    FG_TRY(P.slot.upload(slot.data(), slot.size(), st));
    FG_TRY(P.src.upload(src.data(), src.size(), st));
    std::vector<unsigned short> keep;
    if (L > 0) {
        keep.resize(n);
        for (int q = 0; q < n; ++q) keep[q] = static_cast<unsigned short>(lslot[q] >= 0 ? W + 2 + lslot[q] : W + 1);
        FG_TRY(P.keep.upload(keep.data(), keep.size(), st));
    }
    FG_TRY(P.val64.alloc(slot.size()));
    FG_TRY(P.val32.alloc(slot.size()));
    FG_TRY(P.sink.alloc(kStreamThreads));
    if (upper) {
        FG_TRY(P.dinv64.alloc(n));
        FG_TRY(P.dinv32.alloc(n));
    }
    FG_HIP(hipStreamSynchronize(st));
    P.ok = true;
    return FEMGPU_OK;
}

// Maps between the two streamed schedules (used when both plans exist): for every upper-schedule
// position the lower-schedule position of its row, and for every row its upper-schedule position.
int build_stream_maps(femgpu_ctx *ctx, femgpu_csr *A)
{
    if (!A->slo.ok || !A->sup.ok) return FEMGPU_OK;
    const int n = A->n;
    std::vector<int> rlo(n), rup(n), pos_lo(n), l2u(n), pos_up(n);
    FG_HIP(hipMemcpy(rlo.data(), A->slo.row_of.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    FG_HIP(hipMemcpy(rup.data(), A->sup.row_of.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) {
        pos_lo[rlo[q]] = q;
        pos_up[rup[q]] = q;
    }
    for (int q = 0; q < n; ++q) l2u[q] = pos_lo[rup[q]];
    FG_TRY(A->stream_l2u.upload(l2u.data(), n, ctx->stream));
    FG_TRY(A->stream_pos_up.upload(pos_up.data(), n, ctx->stream));
    FG_TRY(A->stream_t64.alloc(2 * static_cast<size_t>(n)));
    FG_TRY(A->stream_t32.alloc(2 * static_cast<size_t>(n)));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FEMGPU_OK;
}

size_t factor_lds_bytes(const StreamSweep &P)
{
    return sizeof(double) * kFactorRing * (static_cast<size_t>(P.W) + 2) +
           sizeof(int) * (static_cast<size_t>(P.nsteps) + 2 * kStreamPad + 1);
}

// Plan of the streamed factorisation, on the lower sweep's streamed schedule (whose ring slots of
// the L entries are the dependency slots here too). Exists when every row fits the padded format.
int build_stream_factor(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag)
{
    StreamFactor &F = A->sfac;
    F.ok = false;
    const StreamSweep &P = A->slo;
    if (!P.ok || P.K != 4 || P.L != 0 || A->n == 0) return FEMGPU_OK;
    const size_t limit = ctx->lds_limit > 1024 ? ctx->lds_limit - 1024 : ctx->lds_limit;
    if (factor_lds_bytes(P) > limit) return FEMGPU_OK;
    const int n = A->n;
    bool compact = true;
    for (int i = 0; i < n; ++i) {
        const int len = IA[i + 1] - IA[i], nL = diag[i] - IA[i], nU = len - nL - 1;
        if (len > kFactorRow || nL > 4 || nU > kFactorU) return FEMGPU_OK;
        compact = compact && len <= 7 && nL <= 3 && nU <= 3;
    }
    std::vector<int> row_of(n), pos(n);
    FG_HIP(hipMemcpy(row_of.data(), P.row_of.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) pos[row_of[q]] = q;
    // compact rows name each L entry's row by its schedule distance in 11 bits
    for (int q = 0; q < n && compact; ++q) {
        const int i = row_of[q];
        for (int p = IA[i]; p < diag[i]; ++p) compact = compact && q - pos[JA[p]] < 2048;
    }
    std::vector<uint4> codes(compact ? 0 : n);
    std::vector<unsigned long long> rec(compact ? n : 0);
    std::vector<int4> meta(n);
    unsigned seen_valid = 0, seen_none = 0;   // compact: pair positions seen with / without a U column
    for (int q = 0; q < n; ++q) {
        const int i = row_of[q], p0 = IA[i], len = IA[i + 1] - p0, nL = diag[i] - p0;
        unsigned w[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
        unsigned long long r = 0x3fffffffULL;   // compact: 2-bit U columns of the 15 pairs (3: none)
        for (int e = 0; e < len; ++e)
            for (int t = 0; t < nL && t < e; ++t) {
                const int j = JA[p0 + t], col = JA[p0 + e];
                unsigned code = 15u;
                for (int u = 0; u < IA[j + 1] - diag[j] - 1; ++u)
                    if (JA[diag[j] + 1 + u] == col) code = static_cast<unsigned>(u);
                const int bit = 4 * (4 * e + t);
                w[bit / 32] = (w[bit / 32] & ~(15u << (bit % 32))) | (code << (bit % 32));
                if (compact) {   // fixed slots: L entries at 0-2, the diagonal at 3, U at 4-6
                    const int k = factor_pair_index(e < nL ? e : 3 + e - nL, t);
                    if (code < 3u) {
                        r = (r & ~(3ULL << (2 * k))) | (static_cast<unsigned long long>(code) << (2 * k));
                        seen_valid |= 1u << k;
                    } else {
                        seen_none |= 1u << k;
                    }
                }
            }
        if (compact) {
            for (int t = 0; t < nL; ++t)
                r |= static_cast<unsigned long long>(q - pos[JA[p0 + t]]) << (30 + 11 * t);
            rec[q] = r;
        } else {
            // entry 0 has no (entry, L entry) pairs: its 16 code bits carry the row's length and L count
            w[0] = (w[0] & ~0xffu) | static_cast<unsigned>(len) | (static_cast<unsigned>(nL) << 4);
            codes[q] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        meta[q] = make_int4(p0, len, nL, i);
    }
    if (compact) {
        FG_TRY(F.rec.upload(rec.data(), rec.size(), ctx->stream));
    } else {
        FG_TRY(F.codes.upload(codes.data(), codes.size(), ctx->stream));
        FG_TRY(F.dinv.alloc(static_cast<size_t>(n) + kStreamThreads));
    }
    FG_TRY(F.meta.upload(meta.data(), n, ctx->stream));
    FG_TRY(F.aval.alloc(static_cast<size_t>(n) * kFactorRow));
    FG_TRY(F.out.alloc(static_cast<size_t>(n + kStreamThreads) * (compact ? 4 : kFactorRow)));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    F.compact = compact;
    F.structured = compact && (seen_none & kStructAlways) == 0 && (seen_valid & kStructNever) == 0;
    F.ok = true;
    return FEMGPU_OK;
}

int build_factor_plan(femgpu_ctx *ctx, femgpu_csr *A, const int *IA, const int *JA, const int *diag, bool natural,
                      FactorPlan &F)
{
    const int n = A->n;
    std::vector<int> row_of(n);
    if (natural)
        for (int q = 0; q < n; ++q) row_of[q] = q;
    else
        FG_HIP(hipMemcpy(row_of.data(), A->plo.row_of.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int4> info(static_cast<size_t>(n) + 1);
    std::vector<int2> upd;
    upd.reserve(static_cast<size_t>(IA[n]) * 2);
    std::vector<int> slot(n, -1);
    int maxrow = 0, maxupd = 0;
    for (int q = 0; q < n; ++q) {
        const int i = row_of[q], p0 = IA[i], len = IA[i + 1] - p0, nL = diag[i] - p0;
        const size_t u0 = upd.size();
        maxrow = std::max(maxrow, len);
        info[q] = make_int4(p0, len, nL, static_cast<int>(upd.size()));
        for (int p = p0; p < p0 + len; ++p) slot[JA[p]] = p - p0;
        for (int t = 0; t < nL; ++t) {
            const int jrow = JA[p0 + t];
            for (int k = diag[jrow] + 1; k < IA[jrow + 1]; ++k) {
                const int dest = slot[JA[k]];
                if (dest < 0) continue;
                upd.push_back(make_int2(k, dest | (t << 16)));
            }
        }
        for (int p = p0; p < p0 + len; ++p) slot[JA[p]] = -1;
        maxupd = std::max(maxupd, static_cast<int>(upd.size() - u0));
        FG_CHECK(len < 65536 && nL < 32768, FEMGPU_ERR_ARG, "row %d too long for the factorisation plan", i);
    }
    info[n] = make_int4(0, 0, 0, static_cast<int>(upd.size()));   // sentinel: end of the last list
    F.maxrow = maxrow;
    F.nupd = static_cast<int>(upd.size());
    F.maxupd = maxupd;
    FG_TRY(F.rowinfo.upload(info.data(), info.size(), ctx->stream));
    FG_TRY(F.upd.upload(upd.data(), upd.size(), ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FEMGPU_OK;
}

}  // namespace fg
