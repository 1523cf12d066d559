// On-disk formats of the path: the .mesh writer and restart files.
//
// Reference: alexsodt/UMCP
//   surface::saveSurface          src/write_structures.C:473-540
//   Simulation::saveRestart       src/restart.C:7-121
//   Simulation::loadRestart       src/restart.C:123-168
//   surface::writeLimitingSurface src/write_structures.C:662-723
//
// Layouts are the reference's, line for line; floating-point values are
// written with %.17g (the reference's %lf keeps six decimals, which does not
// round-trip FP64).  The readers accept both.
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
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

// surface::writeLimitingSurface (src/write_structures.C:662-723, plim = 1 at
// src/write_structures.C:12): one XYZ frame of the limit positions at the
// three corners (0,0), (0,1), (1,0) of every face, faces in the reference's
// order, positions scaled by alpha as the surface evaluation does
int umcp_write_limit_surface(umcp_surface *s, const double *r, const char *path, int append)
{
    if (!s || !path) return UMCP_ERR_ARG;
    if (!s->prepared) return UMCP_ERR_STATE;
    const Topology &T = s->T;
    const Plan &P = s->P;
    const double *x = r ? r : T.r.data();
    const double al[3] = {r ? r[3 * T.nv] : 1.0, r ? r[3 * T.nv + 1] : 1.0, r ? r[3 * T.nv + 2] : 1.0};
    // (class, index) of every face in reference order
    std::vector<std::pair<int, int>> where((size_t)P.nfaces, {-1, -1});
    for (size_t c = 0; c < P.classes.size(); c++)
        for (int i = 0; i < P.classes[c].nf(); i++) where[P.classes[c].face_id[i]] = {(int)c, i};
    FILE *fp = std::fopen(path, append ? "a" : "w");
    if (!fp) {
        s->err = std::string("cannot write ") + path;
        return UMCP_ERR_ARG;
    }
    std::fprintf(fp, "%d\n", 3 * P.nfaces);
    std::fprintf(fp, "limiting surface\n");
    const int corner_row[3] = {0, 2, 1};   // lattice order (0,0), (0,1), (1,0); corner rows (0,0), (1,0), (0,1)
    for (int f = 0; f < P.nfaces; f++) {
        const FaceClass &C = P.classes[where[f].first];
        const int i = where[f].second, nf = C.nf();
        for (int q = 0; q < 3; q++) {
            const double *w = &C.corner[(size_t)corner_row[q] * C.ncp];
            double p[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < C.ncp; c++) {
                const int v = C.cp[(size_t)c * nf + i];
                const int code = C.img[(size_t)c * nf + i];
                const int ix = code / 9 - 1, iy = (code / 3) % 3 - 1, iz = code % 3 - 1;
                for (int k = 0; k < 3; k++)
                    p[k] += w[c] * (x[3 * v + k] + ix * T.pbc[0][k] + iy * T.pbc[1][k] + iz * T.pbc[2][k]);
            }
            std::fprintf(fp, "C %lf %lf %lf\n", p[0] * al[0], p[1] * al[1], p[2] * al[2]);
        }
    }
    return std::fclose(fp) == 0 ? UMCP_OK : UMCP_ERR_ARG;
}

int umcp_write_limit_psf(umcp_surface *s, const char *path)
{
    if (!s || !path) return UMCP_ERR_ARG;
    if (!s->prepared) return UMCP_ERR_STATE;
    const int nt = s->P.nfaces;
    // three corners per face, lattice order (0,0), (0,1), (1,0); within a
    // face the bonds (0,0)-(1,0), (0,0)-(0,1), (0,1)-(1,0), the order the
    // reference's lattice walk at plim = 1 produces
    FILE *fp = std::fopen(path, "w");
    if (!fp) {
        s->err = std::string("cannot write ") + path;
        return UMCP_ERR_ARG;
    }
    const long natom = 3L * nt, nbond = 3L * nt;
    std::fprintf(fp, "PSF EXT\n\n%8ld !NATOM\n", natom);
    for (long i = 1; i <= natom; i++) {
        char res[32];
        std::snprintf(res, sizeof res, "%ld", i);
        std::fprintf(fp, "%10ld SEGN     %-9sC        C        C        0.000000        1.0000           0   0.00000      0.000000E-01\n",
                     i, res);
    }
    std::fprintf(fp, "%8ld !NBOND: nbonds\n", nbond);
    const int pair[3][2] = {{0, 2}, {0, 1}, {1, 2}};
    long b = 0;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            std::fprintf(fp, "%10ld%10ld", 3L * t + pair[k][0] + 1, 3L * t + pair[k][1] + 1);
            if (++b % 4 == 0) std::fprintf(fp, "\n");
        }
    return std::fclose(fp) == 0 ? UMCP_OK : UMCP_ERR_ARG;
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
