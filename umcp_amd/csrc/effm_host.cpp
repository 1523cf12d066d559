// Setup of the momentum-space integrators: the effective-mass matrix of the
// limit surface (host, FP64, sparse: vertices that share a face).
//
// Reference: alexsodt/UMCP
//   lattice points and mass   src/hd.C:418-481 (plim = 10, 66 points per face)
//   point coefficients        src/uv_map.C:14-200 (get_pt_coeffs)
//   mass matrix               src/modes.C:1719-1743
// (the inverse, dpotrf + dpotri of src/modes.C:1827-1847, and the truncation
// of src/modes.C:1889-1905 run on the device: umcp_effm_build)
//
// The mass matrix sums, over the (plim+1)(plim+2)/2 lattice points
// (i/plim, j/plim) of every face, m c_a c_b for each pair of the point's
// position coefficients.  The lattice is invariant under every permutation of
// the barycentric coordinates, so the sum does not depend on which corner a
// face's parameters start from.  The coefficients of a lattice point depend
// only on the face class (valence), so they are evaluated once per class.
#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "surface_host.hpp"

namespace umcp {

namespace {

constexpr int kPlim = 10;

// position coefficients (the r_w row, cp order) of every lattice point of a
// face class: [npts][ncp]
std::vector<double> lattice_coeffs(const FaceClass &C, int &npts)
{
    const int ncp = C.ncp;
    std::vector<double> out;
    npts = 0;
    std::vector<double> tab(6 * (size_t)ncp);
    for (int fi = 0; fi <= kPlim; fi++)
        for (int fj = 0; fj <= kPlim - fi; fj++, npts++) {
            const double u = fi / (double)kPlim, v = fj / (double)kPlim;
            if (C.valence == 6) {
                box_spline_basis(u, v, tab.data());
            } else if (u < 1e-10 && v < 1e-10) {
                // the extraordinary vertex itself: its limit mask (the owner
                // corner's weights, src/uv_map.C:36-64)
                for (int x = 0; x < ncp; x++) tab[x] = C.corner[x];
            } else {
                irregular_patch_table(C.valence, u, v, tab.data());
            }
            out.insert(out.end(), tab.begin(), tab.begin() + ncp);
        }
    return out;
}

}  // namespace

int effm_points_per_face() { return (kPlim + 1) * (kPlim + 2) / 2; }

int effm_mass_csr(const Topology &T, const Plan &P, double mass_per_pt, std::vector<int> &rowptr,
                  std::vector<int> &col, std::vector<double> &val, std::string &err)
{
    const int nv = T.nv;
    // (column, contribution) per row in accumulation order; merged by a
    // stable sort so each entry sums its contributions in face order
    std::vector<std::vector<std::pair<int, double>>> rows((size_t)nv);
    for (const FaceClass &C : P.classes) {
        const int nf = C.nf(), ncp = C.ncp;
        if (nf == 0) continue;
        if (C.valence != 6 && (int)C.corner.size() < ncp) {
            err = "face class without corner weights";
            return -1;
        }
        int npts = 0;
        const std::vector<double> cf = lattice_coeffs(C, npts);
        // per face: the sum over its lattice points of m c_a c_b for every
        // slot pair (a, b), added once per pair
        std::vector<double> pair((size_t)ncp * ncp);
        std::vector<int> ids(ncp);
        for (int f = 0; f < nf; f++) {
            std::fill(pair.begin(), pair.end(), 0.0);
            for (int p = 0; p < npts; p++) {
                const double *c = &cf[(size_t)p * ncp];
                for (int a = 0; a < ncp; a++) {
                    const double ma = c[a] * mass_per_pt;
                    for (int b = 0; b < ncp; b++) pair[(size_t)a * ncp + b] += ma * c[b];
                }
            }
            for (int x = 0; x < ncp; x++) ids[x] = C.cp[(size_t)x * nf + f];
            for (int a = 0; a < ncp; a++)
                for (int b = 0; b < ncp; b++) rows[(size_t)ids[a]].push_back({ids[b], pair[(size_t)a * ncp + b]});
        }
    }
    rowptr.assign(nv + 1, 0);
    col.clear();
    val.clear();
    for (int i = 0; i < nv; i++) {
        auto &r = rows[(size_t)i];
        std::stable_sort(r.begin(), r.end(), [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
            return x.first < y.first;
        });
        for (size_t k = 0; k < r.size(); k++) {
            if (k > 0 && r[k].first == r[k - 1].first) {
                val.back() += r[k].second;
            } else {
                col.push_back(r[k].first);
                val.push_back(r[k].second);
            }
        }
        rowptr[i + 1] = (int)col.size();
        std::vector<std::pair<int, double>>().swap(r);
    }
    return 0;
}

}  // namespace umcp
