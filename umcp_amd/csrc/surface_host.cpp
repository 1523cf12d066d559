// Setup-time host code: .mesh parsing, ring orientation, face plan, shared
// Gauss-point tables.  Reference behaviour followed, file:line in
// alexsodt/UMCP:
//   parse          src/interpLib.C:1144-1300 (header, optional "sense", vertex lines)
//   orientation    src/interpLib.C:1781-2020 and orientSurface 7515-7600
//   face plan      src/interpLib.C:5436-5553 (ownership by lowest vertex, point_list)
//   image shifts   src/mutil.C:541-599 (MinImage3D), src/interpLib.C:5660-5676
//   basis          src/interpLib.C:5836-5918, 6183-6189 (quartic box splines, ceff_map)
//   irregular      src/interpLib.C:5011-5161 (Atot), 5687-5809 (sub-patch, A^(d-1))
#include "surface_host.hpp"

#include "box_spline.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>

namespace umcp {

#include "quad_tables.inc"

// ---------------------------------------------------------------------------
// parsing
// ---------------------------------------------------------------------------

int parse_mesh(const char *path, Topology &T, std::string &err)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) {
        err = std::string("cannot open mesh file ") + path;
        return -1;
    }
    char line[8192];
    auto next = [&](void) -> bool { return std::fgets(line, sizeof line, fp) != nullptr; };
    if (!next()) { std::fclose(fp); err = "empty mesh file"; return -1; }
    bool three = (line[0] == '3' && (line[1] == 'D' || line[1] == 'd'));
    for (int row = 0; row < (three ? 3 : 2); row++) {
        if (!next()) { std::fclose(fp); err = "truncated periodic box"; return -1; }
        std::sscanf(line, "%lf %lf %lf", &T.pbc[row][0], &T.pbc[row][1], &T.pbc[row][2]);
    }
    if (!three) { T.pbc[2][0] = T.pbc[2][1] = 0.0; T.pbc[2][2] = 1e6; }
    long mark = std::ftell(fp);
    T.fix_sense = 1;
    if (next() && std::strncmp(line, "sense", 5) == 0) std::sscanf(line, "sense %d", &T.fix_sense);
    else std::fseek(fp, mark, SEEK_SET);

    while (next()) {
        if (std::strncmp(line, "ntri", 4) == 0) break;
        if (line[0] == 'P' || line[0] == 'Q') {
            std::fclose(fp);
            err = "protein (P/Q) vertices are not supported on this path";
            return -2;
        }
        int id, val, used = 0;
        double x, y, z;
        if (std::sscanf(line, "%d %lf %lf %lf %d%n", &id, &x, &y, &z, &val, &used) < 5) continue;
        if (val < 5 || val > 7) {  // src/interpLib.C:1287-1291
            std::fclose(fp);
            err = "vertex valence must be between 5 and 7";
            return -3;
        }
        T.r.push_back(x); T.r.push_back(y); T.r.push_back(z);
        T.valence.push_back(val);
        size_t base = T.ring.size();
        T.ring.resize(base + kMaxValence, -1);
        const char *s = line + used;
        for (int e = 0; e < val; e++) {
            int nb, n2 = 0;
            if (std::sscanf(s, "%d%n", &nb, &n2) < 1) {
                std::fclose(fp);
                err = "truncated neighbour list";
                return -4;
            }
            T.ring[base + e] = nb;
            s += n2;
        }
    }
    std::fclose(fp);
    T.nv = (int)T.valence.size();
    if (T.nv == 0) { err = "no vertices"; return -5; }
    for (int i = 0; i < T.nv; i++)
        for (int e = 0; e < T.valence[i]; e++) {
            int j = T.ring[i * kMaxValence + e];
            bool back = false;
            if (j >= 0 && j < T.nv && j != i)
                for (int e2 = 0; e2 < T.valence[j]; e2++) back |= (T.ring[j * kMaxValence + e2] == i);
            if (!back) { err = "neighbour lists are not symmetric"; return -6; }
        }
    return 0;
}

// ---------------------------------------------------------------------------
// orientation
// ---------------------------------------------------------------------------

namespace {

bool adjacent(const Topology &T, int a, int b)
{
    for (int e = 0; e < T.valence[a]; e++)
        if (T.ring[a * kMaxValence + e] == b) return true;
    return false;
}

// First triangle of the reference's triangle enumeration (vertex i ascending,
// j > i in i's ring order, k > j in j's ring order, k adjacent to i).
bool first_triangle(const Topology &T, int tri[3])
{
    for (int i = 0; i < T.nv; i++)
        for (int a = 0; a < T.valence[i]; a++) {
            int j = T.ring[i * kMaxValence + a];
            if (j < i) continue;
            for (int b = 0; b < T.valence[j]; b++) {
                int k = T.ring[j * kMaxValence + b];
                if (k == i || k < j || !adjacent(T, k, i)) continue;
                tri[0] = i; tri[1] = j; tri[2] = k;
                return true;
            }
        }
    return false;
}

// One orientation pass: orient every triangle consistently with `tri0`
// traversed as (t0, t1, t2) when sense > 0, then re-order every ring so that
// (v, ring[e+1], ring[e]) is an oriented triangle, keeping ring[0] first.
int orient_pass(Topology &T, int sense, std::string &err)
{
    int t0[3];
    if (!first_triangle(T, t0)) { err = "mesh has no triangles"; return -1; }
    // successor map per directed edge: oriented triangle (a, b, c) -> succ[a->b] = c
    std::unordered_map<uint64_t, int> succ;      // key (a<<32|b) directed
    auto dkey = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    std::vector<int> stack;
    stack.reserve(64);
    int a0 = t0[0], b0 = sense > 0 ? t0[1] : t0[2], c0 = sense > 0 ? t0[2] : t0[1];
    succ.reserve((size_t)T.nv * 8);
    auto add_tri = [&](int a, int b, int c) -> int {
        uint64_t k1 = dkey(a, b), k2 = dkey(b, c), k3 = dkey(c, a);
        auto it = succ.find(k1);
        if (it != succ.end()) return it->second == c ? 0 : -1;
        if (succ.count(dkey(b, a)) && succ[dkey(b, a)] == c) return -1;  // opposite sense
        succ[k1] = c; succ[k2] = a; succ[k3] = b;
        stack.push_back(a); stack.push_back(b); stack.push_back(c);
        return 1;
    };
    add_tri(a0, b0, c0);
    while (!stack.empty()) {
        int c = stack.back(); stack.pop_back();
        int b = stack.back(); stack.pop_back();
        int a = stack.back(); stack.pop_back();
        int e[3][2] = {{a, b}, {b, c}, {c, a}};
        for (auto &pq : e) {
            int p = pq[0], q = pq[1];
            // the other triangle on edge p-q traverses q -> p; its third vertex
            // is the common neighbour of p and q other than the current one
            int third = (pq[0] == a && pq[1] == b) ? c : (pq[0] == b ? a : b);
            for (int x = 0; x < T.valence[q]; x++) {
                int w = T.ring[q * kMaxValence + x];
                if (w == p || w == third || !adjacent(T, w, p)) continue;
                int rc = add_tri(q, p, w);
                if (rc < 0) { err = "surface is not orientable"; return -2; }
            }
        }
    }
    for (int v = 0; v < T.nv; v++) {
        int val = T.valence[v];
        int tmp[kMaxValence];
        int last = T.ring[v * kMaxValence];
        tmp[0] = last;
        for (int e = 1; e < val; e++) {
            // oriented triangle (v, next, last): succ[last -> v] = next
            auto it = succ.find(dkey(last, v));
            if (it == succ.end()) { err = "could not order the ring of a vertex"; return -3; }
            tmp[e] = last = it->second;
        }
        auto it = succ.find(dkey(last, v));
        if (it == succ.end() || it->second != tmp[0]) { err = "vertex ring is not closed"; return -4; }
        std::memcpy(&T.ring[v * kMaxValence], tmp, sizeof(int) * val);
    }
    return 0;
}

}  // namespace

int orient_mesh(Topology &T, std::string &err)
{
    // loadLattice pass with the file's "sense", then orientSurface with +1
    int rc = orient_pass(T, T.fix_sense, err);
    if (rc == 0) rc = orient_pass(T, 1, err);
    if (rc) return rc;
    T.rev.assign((size_t)T.nv * kMaxValence, -1);
    for (int i = 0; i < T.nv; i++)
        for (int e = 0; e < T.valence[i]; e++) {
            int j = T.ring[i * kMaxValence + e];
            for (int e2 = 0; e2 < T.valence[j]; e2++)
                if (T.ring[j * kMaxValence + e2] == i) T.rev[i * kMaxValence + e] = e2;
        }
    return 0;
}

// ---------------------------------------------------------------------------
// basis tables
// ---------------------------------------------------------------------------

namespace {

struct BoxSpline {
    double coef[12][kMono];  // times 12
};

const BoxSpline &box_spline()
{
    static BoxSpline B = [] {
        BoxSpline b{};
        for (int n = 0; n < 12; n++)
            for (int m = 0; m < kMono; m++) b.coef[n][m] = kBoxTable.c[n][m];
        return b;
    }();
    return B;
}

inline double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

// value and derivatives of u^a v^b w^c along u and v with w = 1 - u - v
void mono_derivs(int a, int b, double u, double v, double out[6])
{
    int c = 4 - a - b;
    double w = 1.0 - u - v;
    auto M = [&](int pa, int pb, int pc) -> double {
        if (pa < 0 || pb < 0 || pc < 0) return 0.0;
        return ipow(u, pa) * ipow(v, pb) * ipow(w, pc);
    };
    out[0] = M(a, b, c);
    // d/du = a u^(a-1) .. - c .. w^(c-1)
    out[1] = a * M(a - 1, b, c) - c * M(a, b, c - 1);
    out[2] = b * M(a, b - 1, c) - c * M(a, b, c - 1);
    out[3] = a * (a - 1) * M(a - 2, b, c) - 2.0 * a * c * M(a - 1, b, c - 1) + c * (c - 1) * M(a, b, c - 2);
    out[4] = a * b * M(a - 1, b - 1, c) - a * c * M(a - 1, b, c - 1) - b * c * M(a, b - 1, c - 1) +
             c * (c - 1) * M(a, b, c - 2);
    out[5] = b * (b - 1) * M(a, b - 2, c) - 2.0 * b * c * M(a, b - 1, c - 1) + c * (c - 1) * M(a, b, c - 2);
}

}  // namespace

void box_spline_basis(double u, double v, double *out)
{
    const BoxSpline &B = box_spline();
    double md[kMono][6];
    for (int a = 0; a <= 4; a++)
        for (int b = 0; a + b <= 4; b++) mono_derivs(a, b, u, v, md[mono_index(a, b)]);
    for (int y = 0; y < 12; y++) {
        const double *cf = B.coef[kSlotBasis[y] - 1];
        for (int d = 0; d < 6; d++) {
            double s = 0.0;
            for (int m = 0; m < kMono; m++)
                if (cf[m] != 0.0) s += cf[m] * md[m][d];
            out[d * 12 + y] = s / 12.0;
        }
    }
}

// ---------------------------------------------------------------------------
// irregular patches
// ---------------------------------------------------------------------------

namespace {

// (N+6) x N Loop stencil for a valence-val patch in point_list order: ring of
// the extraordinary vertex (rows 0..val), the six points around the two other
// corners (rows val+1..val+5) and the six extra points of the next level.
void loop_stencil(int val, std::vector<double> &A)
{
    const int N = val + 6;
    A.assign((size_t)(N + 6) * N, 0.0);
    auto S = [&](int r, int c, double w) { A[(size_t)r * N + c] = w; };
    const double e3 = 3.0 / 8.0, e1 = 1.0 / 8.0, w = 3.0 / (8.0 * val), w6 = 3.0 / 48.0;
    S(0, 0, 1.0 - val * w);
    for (int p = 1; p <= val; p++) S(0, p, w);
    for (int p = 1; p <= val; p++) {  // edge points on the spokes
        int prev = (p == 1) ? val : p - 1, nextp = (p == val) ? 1 : p + 1;
        S(p, 0, e3); S(p, p, e3); S(p, prev, e1); S(p, nextp, e1);
    }
    S(val + 1, 1, e3); S(val + 1, val, e3); S(val + 1, 0, e1); S(val + 1, val + 1, e1);
    for (int c : {0, val, val + 1, val + 2, val + 3, 2}) S(val + 2, c, w6);
    S(val + 2, 1, 1.0 - 6 * w6);
    S(val + 3, 1, e3); S(val + 3, 2, e3); S(val + 3, 0, e1); S(val + 3, val + 3, e1);
    for (int c : {0, 1, val + 3, val + 4, val + 5, 3}) S(val + 4, c, w6);
    S(val + 4, 2, 1.0 - 6 * w6);
    S(val + 5, 2, e3); S(val + 5, 3, e3); S(val + 5, 0, e1); S(val + 5, val + 5, e1);
    const int extra[6][4] = {{1, val + 1, val + 2, val},     {1, val + 2, val + 3, val + 1},
                             {1, val + 3, 2, val + 2},       {2, val + 3, 1, val + 4},
                             {2, val + 4, val + 5, val + 3}, {2, val + 5, 3, val + 4}};
    for (int x = 0; x < 6; x++) {
        S(val + 6 + x, extra[x][0], e3); S(val + 6 + x, extra[x][1], e3);
        S(val + 6 + x, extra[x][2], e1); S(val + 6 + x, extra[x][3], e1);
    }
}

void matmul(int n, const double *A, const double *B, double *C)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int l = 0; l < n; l++) s += A[i * n + l] * B[l * n + j];
            C[i * n + j] = s;
        }
}

}  // namespace

int irregular_patch_table(int val, double fu, double fv, double *out)
{
    if (val < 3 || val > kMaxValence) return -1;
    const int N = val + 6;
    std::vector<double> A;
    loop_stencil(val, A);
    double su = 1.0, sv = 1.0;
    double fd = -std::log10(fu + fv) / std::log10(2.0);
    if (fu + fv >= 1.0 - 1e-9) fd = 1;
    int domain = (int)std::lround(std::ceil(fd));
    double p2 = std::pow(2.0, domain - 1.0);
    su *= p2; sv *= p2; fu *= p2; fv *= p2;
    int pc[12];
    if (fu > 0.5) {
        fv = 2 * fv; fu = 2.0 * fu - 1.0; su *= 2; sv *= 2;
        const int t[12] = {2, val + 3, val + 4, val + 5, 3, 0, 1, val + 2, val + 8, val + 9, val + 10, val + 11};
        std::memcpy(pc, t, sizeof t);
    } else if (fv > 0.5) {
        fu = 2 * fu; fv = 2.0 * fv - 1.0; su *= 2; sv *= 2;
        const int t[12] = {1, val + 2, val + 3, 2, 0, val, val + 1, val + 6, val + 7, val + 8, val + 9, val + 4};
        std::memcpy(pc, t, sizeof t);
    } else {
        fv = 1 - 2 * fv; fu = 1 - 2 * fu; su *= -2; sv *= -2;
        const int t[12] = {val + 3, 2, 1, val + 2, val + 8, val + 9, val + 4, val + 5, 3, 0, val, val + 1};
        std::memcpy(pc, t, sizeof t);
    }
    // P = A^(domain-1) of the top N x N block, by binary powering
    std::vector<double> P((size_t)N * N, 0.0), base((size_t)N * N), tmp((size_t)N * N);
    for (int i = 0; i < N; i++) P[(size_t)i * N + i] = 1.0;
    std::copy(A.begin(), A.begin() + (size_t)N * N, base.begin());
    for (int e = domain - 1; e > 0; e >>= 1) {
        if (e & 1) { matmul(N, P.data(), base.data(), tmp.data()); P.swap(tmp); }
        if (e > 1) { matmul(N, base.data(), base.data(), tmp.data()); base.swap(tmp); }
    }
    double B[6 * 12];
    box_spline_basis(fu, fv, B);
    const double sc[6] = {1.0, su, sv, su * su, su * sv, sv * sv};
    for (int d = 0; d < 6; d++)
        for (int x = 0; x < N; x++) {
            double s = 0.0;
            for (int y = 0; y < 12; y++) {
                // row pc[y] of Atot * P, column x
                double ax = 0.0;
                for (int e = 0; e < N; e++) ax += A[(size_t)pc[y] * N + e] * P[(size_t)e * N + x];
                s += ax * B[d * 12 + y];
            }
            out[d * N + x] = s * sc[d];
        }
    return 0;
}

int irregular_domain_maps(int val, int ndomains, double *out)
{
    if (val < 3 || val > kMaxValence || ndomains < 1) return -1;
    const int N = val + 6;
    std::vector<double> A;
    loop_stencil(val, A);
    // rows of Atot picked by the three sub-patches (the slot tables of
    // irregular_patch_table above)
    const int pcs[3][12] = {
        {2, val + 3, val + 4, val + 5, 3, 0, 1, val + 2, val + 8, val + 9, val + 10, val + 11},
        {1, val + 2, val + 3, 2, 0, val, val + 1, val + 6, val + 7, val + 8, val + 9, val + 4},
        {val + 3, 2, 1, val + 2, val + 8, val + 9, val + 4, val + 5, 3, 0, val, val + 1}};
    std::vector<double> P((size_t)N * N, 0.0), tmp((size_t)N * N);
    for (int i = 0; i < N; i++) P[(size_t)i * N + i] = 1.0;
    for (int d = 1; d <= ndomains; d++) {
        if (d > 1) {  // P = A^(d-1), one more factor of the top N x N block
            matmul(N, A.data(), P.data(), tmp.data());
            P.swap(tmp);
        }
        for (int sub = 0; sub < 3; sub++) {
            double *o = out + ((size_t)(d - 1) * 3 + sub) * 12 * N;
            for (int y = 0; y < 12; y++)
                for (int x = 0; x < N; x++) {
                    double ax = 0.0;
                    for (int e = 0; e < N; e++) ax += A[(size_t)pcs[sub][y] * N + e] * P[(size_t)e * N + x];
                    o[y * N + x] = ax;
                }
        }
    }
    return 0;
}

int irregular_points(int val, double pts[kIrrPts][3])
{
    for (size_t i = 0; i < sizeof(kQuadValences) / sizeof(kQuadValences[0]); i++)
        if (kQuadValences[i] == val) {
            std::memcpy(pts, kQuadPoints[i], sizeof(double) * kIrrPts * 3);
            return 0;
        }
    return -1;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------

namespace {

const double kRegGauss[kRegPts][2] = {{0.1, 0.1}, {0.3, 0.1}, {0.5, 0.1}, {0.7, 0.1}, {0.9, 0.1},
                                      {0.1, 0.3}, {0.3, 0.3}, {0.5, 0.3}, {0.7, 0.3}, {0.1, 0.5},
                                      {0.3, 0.5}, {0.5, 0.5}, {0.1, 0.7}, {0.3, 0.7}, {0.1, 0.9}};

// integer image shift of dr (MinImage3D's greedy 27-neighbour descent)
void image_shift(const double dr_in[3], const double pbc[3][3], int add[3])
{
    double tr[3] = {dr_in[0], dr_in[1], dr_in[2]};
    add[0] = add[1] = add[2] = 0;
    bool moved = true;
    while (moved) {
        moved = false;
        double r2 = tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2];
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    double s[3], t[3];
                    for (int c = 0; c < 3; c++) {
                        s[c] = dx * pbc[0][c] + dy * pbc[1][c] + dz * pbc[2][c];
                        t[c] = tr[c] + s[c];
                    }
                    double n2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
                    if (n2 < r2 - 1e-7) {
                        std::memcpy(tr, t, sizeof tr);
                        r2 = n2;
                        add[0] += dx; add[1] += dy; add[2] += dz;
                        moved = true;
                    }
                }
    }
}

}  // namespace

int build_plan(const Topology &T, Plan &P, std::string &err)
{
    P = Plan();
    // class 0 regular; irregular classes created on first use of a valence
    P.classes.emplace_back();
    {
        FaceClass &R = P.classes[0];
        R.valence = 6; R.ncp = 12; R.npt = kRegPts; R.dudv = 0.5;
        R.table.resize((size_t)kRegPts * 6 * 12);
        R.weight.assign(kRegPts, 1.0 / 15);
        for (int p = 0; p < kRegPts; p++)
            box_spline_basis(kRegGauss[p][0], kRegGauss[p][1], &R.table[(size_t)p * 72]);
        // corners of the volume polyhedron: the box spline at the three corners
        // (src/interpLib.C:426-461 evaluates the volume plan the same way)
        const double uv[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
        R.corner.resize(3 * 12);
        for (int c = 0; c < 3; c++) {
            double B[6 * 12];
            box_spline_basis(uv[c][0], uv[c][1], B);
            for (int y = 0; y < 12; y++) R.corner[c * 12 + y] = B[y];
        }
    }
    struct Face { int cls; int cp[kMaxCp]; uint8_t img[kMaxCp]; bool any; };
    std::vector<Face> faces;
    std::vector<int> irregular_order;  // enumeration order of irregular faces
    int cls_of_val[kMaxValence + 1];
    for (int &c : cls_of_val) c = -1;
    cls_of_val[6] = 0;
    for (int i = 0; i < T.nv; i++) {
        int val = T.valence[i];
        for (int e = 0; e < val; e++) {
            int j = T.nb(i, e), k = T.nb(i, e + 1);
            if (j < i || k < i) continue;
            if (val == 6 && (T.valence[j] != 6 || T.valence[k] != 6)) {
                err = "illegal mesh: a valence-6 face owner touches an irregular vertex "
                      "(subdivide once so irregular vertices own their faces)";
                return -20;
            }
            int ej = T.rev[i * kMaxValence + e], ek = T.rev[i * kMaxValence + (e + 1) % val];
            Face F{};
            int *pl = F.cp;
            pl[0] = i; pl[1] = k; pl[2] = j;
            for (int s = 1; s <= val - 2; s++) pl[2 + s] = T.nb(i, e - s);
            pl[val + 1] = T.nb(k, ek - 2);
            pl[val + 2] = T.nb(k, ek + 3);
            pl[val + 3] = T.nb(k, ek + 2);
            pl[val + 4] = T.nb(j, ej + 3);
            pl[val + 5] = T.nb(j, ej + 2);
            int ncp = val + 6;
            F.any = false;
            for (int c = 0; c < ncp; c++) {
                double dr[3];
                for (int x = 0; x < 3; x++) dr[x] = T.r[3 * pl[c] + x] - T.r[3 * pl[0] + x];
                int add[3];
                image_shift(dr, T.pbc, add);
                for (int x = 0; x < 3; x++)
                    if (add[x] < -1 || add[x] > 1) {
                        err = "control points more than one periodic image apart";
                        return -21;
                    }
                F.img[c] = (uint8_t)((add[0] + 1) * 9 + (add[1] + 1) * 3 + (add[2] + 1));
                F.any |= (F.img[c] != 13);
            }
            if (val == 6) {
                F.cls = 0;
            } else {
                if (cls_of_val[val] < 0) {
                    double pts[kIrrPts][3];
                    if (irregular_points(val, pts)) { err = "no quadrature for this valence"; return -22; }
                    cls_of_val[val] = (int)P.classes.size();
                    P.classes.emplace_back();
                    FaceClass &C = P.classes.back();
                    C.valence = val; C.ncp = ncp; C.npt = kIrrPts; C.dudv = 1.0;
                    C.table.resize((size_t)kIrrPts * 6 * ncp);
                    C.weight.resize(kIrrPts);
                    for (int p = 0; p < kIrrPts; p++) {
                        irregular_patch_table(val, pts[p][0], pts[p][1], &C.table[(size_t)p * 6 * ncp]);
                        C.weight[p] = pts[p][2];
                    }
                    // corners: the owner's limit point is the top eigenvector of
                    // the stencil (src/interpLib.C:455-459), Loop's limit mask
                    // 1/2 on the vertex and 1/(2 val) on its ring (slots 1..val)
                    // for the stencil weight 3/(8 val); the other two corners by
                    // the Stam evaluation
                    C.corner.assign(3 * (size_t)ncp, 0.0);
                    C.corner[0] = 0.5;
                    for (int x = 1; x <= val; x++) C.corner[x] = 0.5 / val;
                    std::vector<double> tb(6 * (size_t)ncp);
                    irregular_patch_table(val, 1.0, 0.0, tb.data());
                    for (int x = 0; x < ncp; x++) C.corner[ncp + x] = tb[x];
                    irregular_patch_table(val, 0.0, 1.0, tb.data());
                    for (int x = 0; x < ncp; x++) C.corner[2 * ncp + x] = tb[x];
                }
                F.cls = cls_of_val[val];
                irregular_order.push_back((int)faces.size());
            }
            faces.push_back(F);
        }
    }
    // global face ids: regular faces first in enumeration order, then irregular
    int nf = (int)faces.size();
    std::vector<int> gid(nf);
    int nreg = 0;
    for (int f = 0; f < nf; f++)
        if (faces[f].cls == 0) gid[f] = nreg++;
    for (size_t x = 0; x < irregular_order.size(); x++) gid[irregular_order[x]] = nreg + (int)x;
    P.nfaces = nf;
    P.nreg = nreg;
    for (int f = 0; f < nf; f++) P.classes[faces[f].cls].face_id.push_back(gid[f]);
    // SoA per class, faces in Morton order of their corner centroids (ties by
    // global id): the 64 faces of a group then form a compact patch whose
    // control points are few distinct vertices (about 95 for 768 (face, slot)
    // incidences on an icosphere), so the face kernel's per-group vertex sums
    // write few gradient records
    std::vector<std::vector<int>> members(P.classes.size());
    for (int f = 0; f < nf; f++) members[faces[f].cls].push_back(f);
    std::vector<uint64_t> key(nf);
    {
        std::vector<double> cen(3 * (size_t)nf);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int f = 0; f < nf; f++)
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int q = 0; q < 3; q++) {
                    const int code = faces[f].img[q];
                    const int ix = code / 9 - 1, iy = (code / 3) % 3 - 1, iz = code % 3 - 1;
                    s += T.r[3 * (size_t)faces[f].cp[q] + k] + ix * T.pbc[0][k] + iy * T.pbc[1][k] + iz * T.pbc[2][k];
                }
                cen[3 * (size_t)f + k] = s / 3.0;
                lo[k] = std::min(lo[k], s / 3.0);
                hi[k] = std::max(hi[k], s / 3.0);
            }
        auto spread = [](uint64_t x) {   // 21 bits -> every third bit
            x &= 0x1fffff;
            x = (x | x << 32) & 0x1f00000000ffffULL;
            x = (x | x << 16) & 0x1f0000ff0000ffULL;
            x = (x | x << 8) & 0x100f00f00f00f00fULL;
            x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
            x = (x | x << 2) & 0x1249249249249249ULL;
            return x;
        };
        for (int f = 0; f < nf; f++) {
            uint64_t m = 0;
            for (int k = 0; k < 3; k++) {
                const double span = hi[k] - lo[k];
                const double t = span > 0 ? (cen[3 * (size_t)f + k] - lo[k]) / span : 0.0;
                m |= spread((uint64_t)(std::min(std::max(t, 0.0), 1.0) * 2097151.0)) << k;
            }
            key[f] = m;
        }
    }
    for (size_t c = 0; c < P.classes.size(); c++) {
        FaceClass &C = P.classes[c];
        auto &mem = members[c];
        std::sort(mem.begin(), mem.end(),
                  [&](int a, int b) { return key[a] != key[b] ? key[a] < key[b] : gid[a] < gid[b]; });
        int n = (int)mem.size();
        C.face_id.resize(n);
        C.cp.resize((size_t)C.ncp * n);
        C.img.resize((size_t)C.ncp * n);
        C.has_img.resize(n);
        for (int x = 0; x < n; x++) {
            const Face &F = faces[mem[x]];
            C.face_id[x] = gid[mem[x]];
            C.has_img[x] = F.any;
            for (int q = 0; q < C.ncp; q++) {
                C.cp[(size_t)q * n + x] = F.cp[q];
                C.img[(size_t)q * n + x] = F.img[q];
            }
        }
    }
    // vertex -> (face, slot) CSR; within a vertex, entries ordered by class,
    // then slot, then face
    std::vector<int> cnt(T.nv + 1, 0);
    for (size_t c = 0; c < P.classes.size(); c++) {
        const FaceClass &C = P.classes[c];
        for (size_t s = 0; s < C.cp.size(); s++) cnt[C.cp[s] + 1]++;
    }
    for (int v = 0; v < T.nv; v++) cnt[v + 1] += cnt[v];
    P.vstart = cnt;
    P.nnz = cnt[T.nv];
    std::vector<int> next(cnt.begin(), cnt.end() - 1);
    for (size_t c = 0; c < P.classes.size(); c++) {
        FaceClass &C = P.classes[c];
        C.pos.resize(C.cp.size());
        for (size_t s = 0; s < C.cp.size(); s++) C.pos[s] = next[C.cp[s]]++;
    }
    // per 64-face group: its distinct control-point vertices (in order of
    // first appearance, slot-major) and each one's incidences slot * 64 + lane
    // in ascending order
    for (size_t c = 0; c < P.classes.size(); c++) {
        FaceClass &C = P.classes[c];
        const int n = C.nf();
        C.gl_off.assign(1, 0);
        C.gl_vert.clear();
        C.gl_inc_off.assign(1, 0);
        C.gl_inc.clear();
        C.loc.assign(C.cp.size(), 0);
        std::vector<int> local(T.nv, -1);
        for (int g0 = 0; g0 < n; g0 += 64) {
            const int g1 = std::min(n, g0 + 64);
            std::vector<int> verts;
            std::vector<std::vector<uint16_t>> inc;
            for (int q = 0; q < C.ncp; q++)
                for (int x = g0; x < g1; x++) {
                    const int v = C.cp[(size_t)q * n + x];
                    if (local[v] < 0) {
                        local[v] = (int)verts.size();
                        verts.push_back(v);
                        inc.emplace_back();
                    }
                    inc[local[v]].push_back((uint16_t)(q * 64 + (x - g0)));
                    C.loc[(size_t)q * n + x] = (uint16_t)local[v];
                }
            for (size_t l = 0; l < verts.size(); l++) {
                C.gl_vert.push_back(verts[l]);
                C.gl_inc.insert(C.gl_inc.end(), inc[l].begin(), inc[l].end());
                C.gl_inc_off.push_back((int)C.gl_inc.size());
                local[verts[l]] = -1;
            }
            C.gl_off.push_back((int)C.gl_vert.size());
        }
    }
    return 0;
}

}  // namespace umcp
