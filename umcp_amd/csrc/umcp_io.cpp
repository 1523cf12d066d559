// On-disk formats of the path: the .mesh writer and restart files.
//
// Reference: alexsodt/UMCP
//   surface::saveSurface          src/write_structures.C:473-540
//   Simulation::saveRestart       src/restart.C:7-121
//   Simulation::loadRestart       src/restart.C:123-168
//
// Layouts are the reference's, line for line; floating-point values are
// written with %.17g (the reference's %lf keeps six decimals, which does not
// round-trip FP64).  The readers accept both.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/umcp.h"
#include "surface_host.hpp"
#include "umcp_state.hpp"

using namespace umcp;

extern "C" {

int umcp_save_mesh(umcp_surface *s, const double *r, const char *path)
{
    if (!s || !path) return UMCP_ERR_ARG;
    const Topology &T = s->T;
    FILE *fp = std::fopen(path, "w");
    if (!fp) {
        s->err = std::string("cannot write ") + path;
        return UMCP_ERR_ARG;
    }
    const double *x = r ? r : T.r.data();
    std::fprintf(fp, "3D saved surface\n");
    for (int row = 0; row < 3; row++)
        std::fprintf(fp, "%.17g %.17g %.17g\n", T.pbc[row][0], T.pbc[row][1], T.pbc[row][2]);
    for (int v = 0; v < T.nv; v++) {
        std::fprintf(fp, "%d %.17g %.17g %.17g %d", v, x[3 * v], x[3 * v + 1], x[3 * v + 2], T.valence[v]);
        for (int e = 0; e < T.valence[v]; e++) std::fprintf(fp, " %d", T.nb(v, e));
        std::fprintf(fp, "\n");
    }
    // triangles once each, from their lowest vertex, in ring order
    std::vector<int> tri;
    for (int i = 0; i < T.nv; i++)
        for (int e = 0; e < T.valence[i]; e++) {
            const int j = T.nb(i, e), k = T.nb(i, e + 1);
            if (i < j && i < k) {
                tri.push_back(i);
                tri.push_back(j < k ? j : k);
                tri.push_back(j < k ? k : j);
            }
        }
    std::fprintf(fp, "ntri %d\n", (int)(tri.size() / 3));
    for (size_t t = 0; t < tri.size(); t += 3) std::fprintf(fp, "%d %d %d\n", tri[t], tri[t + 1], tri[t + 2]);
    const bool ok = std::fclose(fp) == 0;
    return ok ? UMCP_OK : UMCP_ERR_ARG;
}

int umcp_write_restart(const char *path, int nv, const double *r, const double *pp, int seed)
{
    if (!path || !r || nv <= 0) return UMCP_ERR_ARG;
    FILE *fp = std::fopen(path, "w");
    if (!fp) return UMCP_ERR_ARG;
    const double *al = r + 3 * (size_t)nv;
    std::fprintf(fp, "%.17g %.17g %.17g\n", al[0], al[1], al[2]);
    for (int v = 0; v < nv + 1; v++) {
        const double *x = r + 3 * (size_t)v;
        if (pp) {
            // the momenta of the alpha row are zero (the reference allocates
            // 3 nv + 3 momenta and never moves the last three)
            const double p0 = v < nv ? pp[3 * (size_t)v] : 0.0, p1 = v < nv ? pp[3 * (size_t)v + 1] : 0.0,
                         p2 = v < nv ? pp[3 * (size_t)v + 2] : 0.0;
            std::fprintf(fp, "%.17g %.17g %.17g %.17g %.17g %.17g\n", x[0], x[1], x[2], p0, p1, p2);
        } else {
            std::fprintf(fp, "%.17g %.17g %.17g\n", x[0], x[1], x[2]);
        }
    }
    std::fprintf(fp, "seed %d\n", seed);
    return std::fclose(fp) == 0 ? UMCP_OK : UMCP_ERR_ARG;
}

int umcp_read_restart(const char *path, int nv, double *r, double *pp, double *alpha, int *has_pp, int *seed)
{
    if (!path || !r || nv <= 0) return UMCP_ERR_ARG;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return UMCP_ERR_ARG;
    char line[4096];
    double al[3];
    if (!std::fgets(line, sizeof line, fp) || std::sscanf(line, "%lf %lf %lf", al, al + 1, al + 2) != 3) {
        std::fclose(fp);
        return UMCP_ERR_BAD_MESH;
    }
    if (alpha) std::memcpy(alpha, al, sizeof al);
    int six = -1;
    for (int v = 0; v < nv + 1; v++) {
        double x[6];
        if (!std::fgets(line, sizeof line, fp)) {
            std::fclose(fp);
            return UMCP_ERR_BAD_MESH;
        }
        const int n = std::sscanf(line, "%lf %lf %lf %lf %lf %lf", x, x + 1, x + 2, x + 3, x + 4, x + 5);
        if (n != 3 && n != 6) {
            std::fclose(fp);
            return UMCP_ERR_BAD_MESH;
        }
        if (six < 0) six = n == 6;
        if ((n == 6) != (six == 1)) {
            std::fclose(fp);
            return UMCP_ERR_BAD_MESH;
        }
        std::memcpy(r + 3 * (size_t)v, x, 3 * sizeof(double));
        if (pp && n == 6 && v < nv) std::memcpy(pp + 3 * (size_t)v, x + 3, 3 * sizeof(double));
    }
    if (has_pp) *has_pp = six;
    // complexes (not on this path) are skipped; the seed line closes the file
    int sd = -1;
    while (std::fgets(line, sizeof line, fp))
        if (std::strncmp(line, "seed", 4) == 0) std::sscanf(line, "seed %d", &sd);
    if (seed) *seed = sd;
    std::fclose(fp);
    return UMCP_OK;
}

}  // extern "C"
