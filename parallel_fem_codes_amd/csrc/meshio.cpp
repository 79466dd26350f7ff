// Mesh generation and mesh files (host code): the structured generator of
// Mesh_Generator/Structured_Mesh/meshGenerator.c (node and element numbering, boundary marks and
// their Type table), the generator's output file `<PROBLEM>_<nnodes>_<nel>.dat`, and the per-rank
// file `RANK_<r>_<PROBLEM>_<nnodes>_<nel>.dat` that Parallel_Preprocess writes and the drivers'
// Preprocess.c reads (01_SS_Transport_Equation/2D/01_CommonFiles/Preprocess.c:127-168).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

using namespace fg;

namespace {

// Type of each of the 4 possible dofs for a boundary mark (meshGenerator.c:16-19): mark m has
// Type 0 (prescribed) for dof j when bit (3 - j) of m is set.
int mark_type(int mark, int j) { return ((mark >> (3 - j)) & 1) ? 0 : 1; }

bool ieq(const char *a, const char *b)
{
    for (; *a && *b; ++a, ++b)
        if (std::tolower(static_cast<unsigned char>(*a)) != std::tolower(static_cast<unsigned char>(*b))) return false;
    return *a == *b;
}

struct Reader {
    FILE *f = nullptr;
    explicit Reader(const char *path) : f(std::fopen(path, "r")) {}
    ~Reader() {
        if (f) std::fclose(f);
    }
    bool num(int &v) { return std::fscanf(f, "%d", &v) == 1; }
    bool num(double &v) { return std::fscanf(f, "%lf", &v) == 1; }
};

}  // namespace

extern "C" int femgpu_structured_mesh(int M, int N, double xinf, double yinf, double xsup, double ysup, int right,
                                      double *x, double *y, int *ien)
{
    FG_CHECK(M >= 1 && N >= 1 && x && y && ien && xsup > xinf && ysup > yinf, FEMGPU_ERR_ARG,
             "femgpu_structured_mesh: bad arguments");
    const int nosx = M + 1, nosy = N + 1, nel = 2 * M * N;
    const double hx = (xsup - xinf) / M, hy = (ysup - yinf) / N;
    // elements: two per square, squares row by row, the cut chosen by `right`
    // (meshGenerator.c:113-179; 0-based vertex numbers of the counters there)
    for (int e = 0; e < nel; ++e) {
        const int sq = e / 2, row = sq / M, col = sq % M;
        const int a = row * nosx + col;   // lower-left node of the square
        int *t = ien + 3 * static_cast<size_t>(e);
        if (right) {
            if (e % 2 == 0) { t[0] = a; t[1] = a + 1; t[2] = a + 1 + nosx; }
            else            { t[0] = a; t[1] = a + 1 + nosx; t[2] = a + nosx; }
        } else {
            if (e % 2 == 0) { t[0] = a; t[1] = a + 1; t[2] = a + nosx; }
            else            { t[0] = a + 1; t[1] = a + 1 + nosx; t[2] = a + nosx; }
        }
    }
    // nodes row by row (:182-189)
    for (int j = 0, k = 0; j < nosy; ++j)
        for (int i = 0; i < nosx; ++i, ++k) {
            x[k] = xinf + i * hx;
            y[k] = yinf + j * hy;
        }
    return FEMGPU_OK;
}

extern "C" int femgpu_structured_marks(const char *problem, int nnodes, const double *x, const double *y, int *mark,
                                       int *type)
{
    FG_CHECK(problem && nnodes >= 0 && x && y && mark, FEMGPU_ERR_ARG, "femgpu_structured_marks: bad arguments");
    auto on = [](double v, double c) { return std::fabs(v - c) <= 1e-15; };
    int ndof = 0;
    // setmark_according_boundary (meshGenerator.c:224-360)
    if (ieq(problem, "OBLIQUO") || ieq(problem, "REFLETIDO")) {
        ndof = 4;
        for (int i = 0; i < nnodes; ++i)
            mark[i] = on(y[i], 0.0) ? 2 : (on(x[i], 0.0) || on(y[i], 1.0)) ? 15 : 0;
    } else if (ieq(problem, "EXPLOSION")) {
        ndof = 4;
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 0.0) || on(y[i], 0.0) || on(x[i], 2.0) || on(y[i], 2.0)) ? 15 : 0;
    } else if (ieq(problem, "SOD")) {
        ndof = 4;
        double ymax = nnodes ? y[0] : 0.0;
        for (int i = 0; i < nnodes; ++i) ymax = std::max(ymax, y[i]);
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 0.0) || on(x[i], 1.0)) ? 15 : (on(y[i], 0.0) || on(y[i], ymax)) ? 2 : 0;
    } else if (ieq(problem, "BAROCLINIC")) {
        ndof = 4;
        for (int i = 0; i < nnodes; ++i) mark[i] = (on(y[i], 0.0) || on(y[i], 8.0)) ? 15 : 0;
    } else if (ieq(problem, "PUDIM")) {
        ndof = 1;
        // the reference's last test reads fabs(coord[i][1] <= 0), i.e. the truth value of y <= 0
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 1.0) || on(x[i], -1.0) || on(y[i], 1.0) || on(y[i], -1.0) ||
                       (std::fabs(x[i]) <= 1e-15 && y[i] <= 0.0))
                          ? 15
                          : 0;
    } else if (ieq(problem, "CONE")) {
        ndof = 1;
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 10.0) || on(x[i], 0.0) || on(y[i], 10.0) || on(y[i], 0.0)) ? 15 : 0;
    } else if (ieq(problem, "CARTOLA") || ieq(problem, "TESTE")) {
        ndof = 1;
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 1.0) || on(x[i], 0.0) || on(y[i], 1.0) || on(y[i], 0.0)) ? 15 : 0;
    } else if (ieq(problem, "CAVITY")) {
        ndof = 3;
        for (int i = 0; i < nnodes; ++i)
            mark[i] = (on(x[i], 1.0) || on(x[i], 0.0) || on(y[i], 1.0) || on(y[i], 0.0)) ? 12
                      : (on(x[i], 0.5) && on(y[i], 0.5))                                  ? 2
                                                                                          : 0;
    } else {
        set_error("femgpu_structured_marks: problem \"%s\" not defined", problem);
        return FEMGPU_ERR_ARG;
    }
    if (type)
        for (int i = 0; i < nnodes; ++i)
            for (int j = 0; j < ndof; ++j) type[static_cast<size_t>(i) * ndof + j] = mark_type(mark[i], j);
    return ndof;
}

// Mesh file, format 0 (generator output, meshGenerator.c:193-211): nnodes, then per node
// "x y" and ndof Types; nel, then per element three vertices (and -1 when ndof > 1).
// Format 1 (one rank's file, Preprocess.c:127-168 / Parallel_Preprocess/Preprocess.c:270-318):
// nnodes, then per node "x y" and per dof either "0" or "Type Send invP_id"; nel, then per element
// three vertices (and the element Type when ndof > 1).
// Call with null arrays to learn the sizes (info), then with arrays of those sizes.
extern "C" int femgpu_mesh_file_read(const char *path, int format, int ndof, femgpu_mesh_file_info *info, double *x,
                                     double *y, int *type, int *invP, int *ien, int *etype)
{
    FG_CHECK(path && info && (format == 0 || format == 1) && ndof >= 1 && ndof <= 4, FEMGPU_ERR_ARG,
             "femgpu_mesh_file_read: bad arguments");
    Reader r(path);
    FG_CHECK(r.f, FEMGPU_ERR_ARG, "femgpu_mesh_file_read: cannot open %s", path);
    int nn = 0;
    FG_CHECK(r.num(nn) && nn >= 0, FEMGPU_ERR_ARG, "%s: bad node count", path);
    const bool fill = x && y && type;
    for (int i = 0; i < nn; ++i) {
        double xi, yi;
        FG_CHECK(r.num(xi) && r.num(yi), FEMGPU_ERR_ARG, "%s: node %d: bad coordinates", path, i);
        if (fill) {
            x[i] = xi;
            y[i] = yi;
        }
        for (int j = 0; j < ndof; ++j) {
            int t = 0, send = 0, id = -1;
            FG_CHECK(r.num(t), FEMGPU_ERR_ARG, "%s: node %d: bad Type", path, i);
            if (format == 1 && t != 0)
                FG_CHECK(r.num(send) && r.num(id), FEMGPU_ERR_ARG, "%s: node %d: bad Send/invP_id", path, i);
            if (fill) type[static_cast<size_t>(i) * ndof + j] = t;
            if (invP) invP[static_cast<size_t>(i) * ndof + j] = id;
        }
    }
    int ne = 0;
    FG_CHECK(r.num(ne) && ne >= 0, FEMGPU_ERR_ARG, "%s: bad element count", path);
    for (int e = 0; e < ne; ++e) {
        int v[3], et = -1;
        FG_CHECK(r.num(v[0]) && r.num(v[1]) && r.num(v[2]), FEMGPU_ERR_ARG, "%s: element %d: bad vertices", path, e);
        if (ndof > 1) FG_CHECK(r.num(et), FEMGPU_ERR_ARG, "%s: element %d: bad element Type", path, e);
        for (int k = 0; k < 3; ++k)
            FG_CHECK(v[k] >= 0 && v[k] < nn, FEMGPU_ERR_ARG, "%s: element %d: vertex %d out of range", path, e, v[k]);
        if (ien)
            for (int k = 0; k < 3; ++k) ien[3 * static_cast<size_t>(e) + k] = v[k];
        if (etype) etype[e] = et;
    }
    info->nnodes = nn;
    info->nel = ne;
    info->ndof = ndof;
    info->format = format;
    return FEMGPU_OK;
}

// Writes format 0 exactly as meshGenerator.c:193-211 prints it.
extern "C" int femgpu_mesh_file_write(const char *path, int nnodes, int nel, int ndof, const double *x, const double *y,
                                      const int *type, const int *ien)
{
    FG_CHECK(path && nnodes >= 0 && nel >= 0 && ndof >= 1 && ndof <= 4 && x && y && type && ien, FEMGPU_ERR_ARG,
             "femgpu_mesh_file_write: bad arguments");
    FILE *f = std::fopen(path, "w");
    FG_CHECK(f, FEMGPU_ERR_ARG, "femgpu_mesh_file_write: cannot open %s", path);
    std::fprintf(f, "%d\n", nnodes);
    for (int i = 0; i < nnodes; ++i) {
        std::fprintf(f, "%.14lf\t%.14lf\t", x[i], y[i]);
        for (int j = 0; j < ndof; ++j) std::fprintf(f, "%d\t", type[static_cast<size_t>(i) * ndof + j]);
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "%d\n", nel);
    for (int e = 0; e < nel; ++e) {
        const int *t = ien + 3 * static_cast<size_t>(e);
        if (ndof == 1) std::fprintf(f, "%d \t %d \t %d\n", t[0], t[1], t[2]);
        else std::fprintf(f, "%d \t %d \t %d \t -1\n", t[0], t[1], t[2]);
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    FG_CHECK(ok, FEMGPU_ERR_ARG, "femgpu_mesh_file_write: write to %s failed", path);
    return FEMGPU_OK;
}
