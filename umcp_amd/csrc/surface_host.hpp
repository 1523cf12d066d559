// Host-side model of one membrane surface: mesh topology in the reference's
// orientation, the face plan and the shared Gauss-point tables.  Everything
// here runs once per mesh (setup); the per-step work is on the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace umcp {

constexpr int kMaxValence = 12;
constexpr int kMaxCp = kMaxValence + 6;
constexpr int kRegPts = 15;  // regular-face Gauss points (src/interpLib.C:5336-5375)
constexpr int kIrrPts = 16;  // 4 x 4 irregular-face points (src/interpLib.C:5608-5645)

// Vertex ring structure as loadLattice leaves it (src/interpLib.C:1144-2606).
struct Topology {
    int nv = 0;
    double pbc[3][3] = {{1e10, 0, 0}, {0, 1e10, 0}, {0, 0, 1e10}};
    int fix_sense = 1;
    std::vector<double> r;      // 3 nv, positions from the file
    std::vector<int> valence;   // nv
    std::vector<int> ring;      // nv * kMaxValence, oriented neighbour rings
    std::vector<int> rev;       // position of i in the ring of ring[i][e]
    int nb(int v, int e) const {
        int n = valence[v];
        return ring[v * kMaxValence + ((e % n) + n) % n];
    }
};

// All faces with the same valence at their owning vertex share one table set.
struct FaceClass {
    int valence = 6;
    int ncp = 12;
    int npt = kRegPts;
    double dudv = 0.5;                 // regular faces integrate over half the unit square
    std::vector<int> face_id;          // global face index (reference face order)
    std::vector<int> cp;               // [ncp][nf]   control-point vertex ids (SoA)
    std::vector<uint8_t> img;          // [ncp][nf]   periodic image code, 13 = none
    std::vector<uint8_t> has_img;      // [nf]        any image shift on this face
    std::vector<double> table;         // [npt][6][ncp] r_w, r_u, r_v, r_uu, r_uv, r_vv
    std::vector<double> weight;        // [npt]
    // position weights of the three corners (u, v) = (0,0), (1,0), (0,1): the
    // limit points of the owner i, of ring[e] and of ring[e+1]  [3][ncp]
    std::vector<double> corner;
    std::vector<int> pos;              // [ncp][nf]   CSR position of (face, slot)
    // Per group of 64 faces (faces 64 g .. 64 g + 63): the distinct vertices
    // among its control points, gl_vert[gl_off[g] .. gl_off[g + 1]), and for
    // the class-wide local vertex l its incidences gl_inc[gl_inc_off[l] ..
    // gl_inc_off[l + 1]) as slot * 64 + (face - 64 g), ascending.  The face
    // kernel writes one gradient record per local vertex (the sum over its
    // incidences in that order).
    std::vector<int> gl_off;           // ngroups + 1
    std::vector<int> gl_vert;          // local vertices
    std::vector<int> gl_inc_off;       // local vertices + 1
    std::vector<uint16_t> gl_inc;      // ncp * nf
    std::vector<uint16_t> loc;         // [ncp][nf] index of (slot, face)'s vertex in its group's list
    int nf() const { return (int)face_id.size(); }
};

struct Plan {
    int nfaces = 0;
    int nreg = 0;
    std::vector<FaceClass> classes;    // classes[0] is the regular class
    // vertex -> (face, slot) incidence as CSR, entries of a vertex ordered by
    // class, slot, face: entry pos[c][f] is slot c of face f
    std::vector<int> vstart;           // nv + 1
    int nnz = 0;
};

int parse_mesh(const char *path, Topology &T, std::string &err);
int orient_mesh(Topology &T, std::string &err);
int build_plan(const Topology &T, Plan &P, std::string &err);

// effective-mass matrix of the momentum-space integrators (effm_host.cpp):
// M as CSR, rows ascending, columns ascending
int effm_points_per_face();
int effm_mass_csr(const Topology &T, const Plan &P, double mass_per_pt, std::vector<int> &rowptr,
                  std::vector<int> &col, std::vector<double> &val, std::string &err);
// the Chebyshev solve of M x = b (umcp_effm_build_sparse): a reverse
// Cuthill-McKee order of the vertices; M in that order in ELL layout
// (slot-major [width][n], columns ascending, rows padded with zeros on the
// diagonal) and D^-1;
// the spectral interval [lo, hi] of D^-1 M (0 on success); the iterations
// that bring the residual below tol on it; the per-iteration coefficients
// (rows and columns in the order perm: new index p is vertex perm[p])
std::vector<int> effm_rcm_order(const std::vector<int> &rowptr, const std::vector<int> &col);
void effm_ell(const std::vector<int> &rowptr, const std::vector<int> &col, const std::vector<double> &val,
              const std::vector<int> &perm, int &width, std::vector<int> &ell_col, std::vector<double> &ell_val,
              std::vector<double> &dinv);
int effm_jacobi_interval(const std::vector<int> &rowptr, const std::vector<int> &col, const std::vector<double> &val,
                         double &lo, double &hi);
int effm_cheb_iterations(double lo, double hi, double tol);
void effm_cheb_coeffs(double lo, double hi, int iters, std::vector<double> &c1, std::vector<double> &c2,
                      double &inv_theta);

// 12 quartic box-spline basis functions and their first/second derivatives at
// (u, v), in the control-point slot order of the face plan: out[d * 12 + y].
void box_spline_basis(double u, double v, double *out);
// Stam evaluation of an irregular patch of valence val at (u, v): out[d*(val+6)+x].
int irregular_patch_table(int val, double u, double v, double *out);
// Sub-patch maps of a valence-val irregular face for levels 1..ndomains, the
// tables irr_kernel::setup precomputes (src/irr_kernel.C:10-260):
// out[((d - 1) * 3 + sub) * 12 * (val + 6) + y * (val + 6) + x] is the weight
// of control point x in box-spline slot y of sub-patch sub at level d, i.e.
// row pc_sub[y] of Atot A^(d-1).
int irregular_domain_maps(int val, int ndomains, double *out);
// The 16 (u, v, weight) points of an irregular face of valence val.
int irregular_points(int val, double pts[kIrrPts][3]);

}  // namespace umcp
