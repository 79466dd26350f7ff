// Delaunay triangulation of a planar point set, for the serving workload's unstructured meshes (the
// reference builds unstructured meshes with gmsh through MG/Unstructured_Mesh/gmsh2Mesh.c; gmsh is
// not part of the reference tree). Bowyer-Watson insertion with ghost triangles: every hull edge
// (a, b) has a ghost (b, a, inf) outside it, so points outside the current hull are inserted by the
// same cavity rule as interior ones, and collinear hull points stay vertices.
//   * insertion order: Hilbert curve over a 2^16 grid (spatially coherent, so the walk from the
//     previous insertion is short);
//   * location: visibility walk over the real triangles;
//   * cavity: the triangles whose circumcircle holds the point strictly (real), or whose outer
//     half-plane holds it strictly or whose hull edge holds it strictly inside (ghost);
//   * predicates in long double (x86 80-bit) on the input coordinates.
// Duplicate points are skipped (not vertices of any triangle) and counted.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "common.h"

namespace {

constexpr int kInf = -1;

struct Tri {
    int v[3];    // counter-clockwise; a ghost has v[2] == kInf and its real edge v[0] -> v[1]
    int nb[3];   // nb[i]: the triangle across the edge opposite v[i] (v[i+1] -> v[i+2])
    bool alive;
};

struct Delaunay {
    const double *x, *y;
    std::vector<Tri> T;
    std::vector<int> free_;
    int last = 0;

    long double orient(int a, int b, int c) const
    {
        const long double ax = x[a], ay = y[a];
        return (static_cast<long double>(x[b]) - ax) * (static_cast<long double>(y[c]) - ay) -
               (static_cast<long double>(y[b]) - ay) * (static_cast<long double>(x[c]) - ax);
    }
    // > 0: d strictly inside the circumcircle of the counter-clockwise triangle (a, b, c)
    long double incircle(int a, int b, int c, int d) const
    {
        const long double dx = x[d], dy = y[d];
        const long double adx = x[a] - dx, ady = y[a] - dy, bdx = x[b] - dx, bdy = y[b] - dy;
        const long double cdx = x[c] - dx, cdy = y[c] - dy;
        const long double al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
        return adx * (bdy * cl - cdy * bl) - ady * (bdx * cl - cdx * bl) + al * (bdx * cdy - bdy * cdx);
    }
    bool ghost(int t) const { return T[t].v[2] == kInf; }
    // p strictly inside segment (a, b), given collinear
    bool on_segment(int a, int b, int p) const
    {
        const long double ux = static_cast<long double>(x[b]) - x[a], uy = static_cast<long double>(y[b]) - y[a];
        const long double s = (static_cast<long double>(x[p]) - x[a]) * ux + (static_cast<long double>(y[p]) - y[a]) * uy;
        return s > 0 && s < ux * ux + uy * uy;
    }
    bool conflict(int t, int p) const
    {
        const Tri &r = T[t];
        if (r.v[2] == kInf) {
            const long double o = orient(r.v[0], r.v[1], p);
            return o > 0 || (o == 0 && on_segment(r.v[0], r.v[1], p));
        }
        return incircle(r.v[0], r.v[1], r.v[2], p) > 0;
    }
    int alloc()
    {
        if (!free_.empty()) {
            const int t = free_.back();
            free_.pop_back();
            return t;
        }
        T.push_back(Tri{});
        return static_cast<int>(T.size()) - 1;
    }
    int make(int a, int b, int c)
    {
        // ghosts keep kInf in slot 2 (rotation preserves orientation)
        if (a == kInf) return make(b, c, a);
        if (b == kInf) return make(c, a, b);
        const int t = alloc();
        T[t] = Tri{{a, b, c}, {-1, -1, -1}, true};
        return t;
    }
    static int edge_index(const Tri &r, int a, int b)   // the i with v[i+1] = a, v[i+2] = b
    {
        for (int i = 0; i < 3; ++i)
            if (r.v[(i + 1) % 3] == a && r.v[(i + 2) % 3] == b) return i;
        return -1;
    }

    // A real triangle whose closure holds p, or a ghost in conflict with p (p outside the hull).
    // Returns -1 when p repeats a vertex.
    int locate(int p, int &dup)
    {
        dup = 0;
        int t = last;
        if (t < 0 || t >= static_cast<int>(T.size()) || !T[t].alive || ghost(t)) {
            t = -1;
            for (int k = 0; k < static_cast<int>(T.size()); ++k)
                if (T[k].alive && !ghost(k)) {
                    t = k;
                    break;
                }
        }
        const long long limit = 4LL * static_cast<long long>(T.size()) + 64;
        unsigned rot = static_cast<unsigned>(p);
        for (long long step = 0; step < limit; ++step) {
            const Tri &r = T[t];
            int next = -1;
            const int s = static_cast<int>(rot++ % 3u);
            for (int k = 0; k < 3 && next < 0; ++k) {
                const int i = (s + k) % 3;
                if (orient(r.v[(i + 1) % 3], r.v[(i + 2) % 3], p) < 0) next = r.nb[i];
            }
            if (next < 0) {
                for (int i = 0; i < 3; ++i)
                    if (x[r.v[i]] == x[p] && y[r.v[i]] == y[p]) {
                        dup = 1;
                        return -1;
                    }
                return t;
            }
            if (ghost(next)) return next;
            t = next;
        }
        // the walk did not settle (round-off): scan
        for (int k = 0; k < static_cast<int>(T.size()); ++k) {
            if (!T[k].alive) continue;
            if (ghost(k)) {
                if (conflict(k, p)) return k;
                continue;
            }
            const Tri &r = T[k];
            if (orient(r.v[0], r.v[1], p) >= 0 && orient(r.v[1], r.v[2], p) >= 0 && orient(r.v[2], r.v[0], p) >= 0)
                return k;
        }
        return -2;
    }

    int insert(int p)
    {
        int dup = 0;
        int t0 = locate(p, dup);
        if (dup) return 1;
        if (t0 < 0) return -1;
        if (!conflict(t0, p)) {   // p on an edge of t0 shared with a ghost: start from the ghost
            for (int i = 0; i < 3; ++i)
                if (T[t0].nb[i] >= 0 && conflict(T[t0].nb[i], p)) {
                    t0 = T[t0].nb[i];
                    break;
                }
            if (!conflict(t0, p)) return -1;
        }
        // cavity by breadth-first search over conflicting neighbours
        std::vector<int> cav{t0}, bnd_t, bnd_i;
        T[t0].alive = false;
        for (size_t q = 0; q < cav.size(); ++q) {
            const int t = cav[q];
            for (int i = 0; i < 3; ++i) {
                const int u = T[t].nb[i];
                if (u >= 0 && T[u].alive && conflict(u, p)) {
                    T[u].alive = false;
                    cav.push_back(u);
                } else if (u < 0 || T[u].alive) {
                    bnd_t.push_back(t);
                    bnd_i.push_back(i);
                }
            }
        }
        // new fan: one triangle per cavity boundary edge (a, b) + p
        struct Edge {
            int a, b, out, nt;
        };
        std::vector<Edge> E(bnd_t.size());
        for (size_t k = 0; k < bnd_t.size(); ++k) {
            const Tri &r = T[bnd_t[k]];
            const int i = bnd_i[k];
            E[k] = {r.v[(i + 1) % 3], r.v[(i + 2) % 3], r.nb[i], -1};
        }
        for (int t : cav) free_.push_back(t);
        for (auto &e : E) {
            e.nt = make(e.a, e.b, p);
            Tri &r = T[e.nt];
            const int i = edge_index(r, e.a, e.b);
            r.nb[i] = e.out;
            if (e.out >= 0) {
                Tri &o = T[e.out];
                const int j = edge_index(o, e.b, e.a);
                o.nb[j] = e.nt;
            }
        }
        // link the fan: the triangle with edge (b, p) meets the one with edge (p, b)
        for (auto &e : E) {
            Tri &r = T[e.nt];
            const int ib = edge_index(r, e.b, p);
            for (auto &f : E)
                if (f.a == e.b) {
                    r.nb[ib] = f.nt;
                    break;
                }
            const int ia = edge_index(r, p, e.a);
            for (auto &f : E)
                if (f.b == e.a) {
                    r.nb[ia] = f.nt;
                    break;
                }
            if (!ghost(e.nt)) last = e.nt;
        }
        return 0;
    }
};

uint32_t hilbert(uint32_t xx, uint32_t yy, int order)
{
    uint32_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        const uint32_t rx = (xx & s) ? 1 : 0, ry = (yy & s) ? 1 : 0;
        d += s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                xx = s - 1 - xx;
                yy = s - 1 - yy;
            }
            std::swap(xx, yy);
        }
    }
    return d;
}

}  // namespace

extern "C" int femgpu_delaunay(int n, const double *x, const double *y, int max_tri, int *ntri, int *tri,
                               int *duplicates)
{
    FG_CHECK(n >= 3 && x && y && ntri && tri && max_tri >= 0, FEMGPU_ERR_ARG, "femgpu_delaunay: bad arguments");
    for (int i = 0; i < n; ++i)
        FG_CHECK(std::isfinite(x[i]) && std::isfinite(y[i]), FEMGPU_ERR_ARG, "femgpu_delaunay: point %d not finite", i);
    double x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0];
    for (int i = 1; i < n; ++i) {
        x0 = std::min(x0, x[i]);
        x1 = std::max(x1, x[i]);
        y0 = std::min(y0, y[i]);
        y1 = std::max(y1, y[i]);
    }
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<uint32_t> key(n);
    for (int i = 0; i < n; ++i)
        key[i] = hilbert(static_cast<uint32_t>((x[i] - x0) * sx), static_cast<uint32_t>((y[i] - y0) * sy), 16);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });

    Delaunay D;
    D.x = x;
    D.y = y;
    // initial triangle: the first point, the next distinct one, the first not collinear with them
    const int a = order[0];
    int b = -1, ib = -1;
    for (int k = 1; k < n && b < 0; ++k)
        if (x[order[k]] != x[a] || y[order[k]] != y[a]) {
            b = order[k];
            ib = k;
        }
    FG_CHECK(b >= 0, FEMGPU_ERR_ARG, "femgpu_delaunay: all points coincide");
    int c = -1, ic = -1;
    for (int k = ib + 1; k < n && c < 0; ++k)
        if (D.orient(a, b, order[k]) != 0) {
            c = order[k];
            ic = k;
        }
    FG_CHECK(c >= 0, FEMGPU_ERR_ARG, "femgpu_delaunay: all points are collinear");
    if (D.orient(a, b, c) < 0) std::swap(b, c);
    D.T.reserve(2 * static_cast<size_t>(n) + 16);
    const int t = D.make(a, b, c);
    const int g0 = D.make(b, a, kInf), g1 = D.make(c, b, kInf), g2 = D.make(a, c, kInf);
    // t: nb[0] across (b, c), nb[1] across (c, a), nb[2] across (a, b)
    D.T[t].nb[0] = g1;
    D.T[t].nb[1] = g2;
    D.T[t].nb[2] = g0;
    // ghost (u, v, inf): nb[2] across (u, v) is the real triangle; nb[0] across (v, inf), nb[1] across (inf, u)
    auto link_ghost = [&](int g, int real, int next_g, int prev_g) {
        D.T[g].nb[2] = real;
        D.T[g].nb[0] = next_g;
        D.T[g].nb[1] = prev_g;
    };
    // hull a -> b -> c -> a (counter-clockwise); ghost over edge (a,b) is (b,a,inf) = g0.
    // (v, inf) of g0 = (a, inf) meets ghost (a, c, inf) = g2's (inf, a); (inf, u) of g0 = (inf, b)
    // meets g1 = (c, b, inf)'s (b, inf).
    link_ghost(g0, t, g2, g1);
    link_ghost(g1, t, g0, g2);
    link_ghost(g2, t, g1, g0);
    D.last = t;
    int dups = 0;
    for (int k = 0; k < n; ++k) {
        const int p = order[k];
        if (k == 0 || k == ib || k == ic) continue;
        const int rc = D.insert(p);
        FG_CHECK(rc >= 0, FEMGPU_ERR_STATE, "femgpu_delaunay: point %d could not be inserted", p);
        dups += rc;
    }
    int m = 0;
    for (const Tri &r : D.T)
        if (r.alive && r.v[2] != kInf) ++m;
    FG_CHECK(m <= max_tri, FEMGPU_ERR_ARG, "femgpu_delaunay: %d triangles exceed max_tri = %d", m, max_tri);
    int k = 0;
    for (const Tri &r : D.T)
        if (r.alive && r.v[2] != kInf) {
            tri[3 * k] = r.v[0];
            tri[3 * k + 1] = r.v[1];
            tri[3 * k + 2] = r.v[2];
            ++k;
        }
    *ntri = m;
    if (duplicates) *duplicates = dups;
    return FEMGPU_OK;
}
