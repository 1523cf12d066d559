// Setup of the momentum-space integrators: the effective-mass matrix of the
// limit surface and its truncated inverse (host, FP64, O(nv^3)).
//
// Reference: alexsodt/UMCP
//   lattice points and mass   src/hd.C:418-481 (plim = 10, 66 points per face)
//   point coefficients        src/uv_map.C:14-200 (get_pt_coeffs)
//   mass matrix and inverse   src/modes.C:1719-1847 (dpotrf + dpotri)
//   truncation                src/modes.C:16 (rel_tol), 1889-1905
//
// The mass matrix sums, over the (plim+1)(plim+2)/2 lattice points
// (i/plim, j/plim) of every face, m c_a c_b for each pair of the point's
// position coefficients.  The lattice is invariant under every permutation of
// the barycentric coordinates, so the sum does not depend on which corner a
// face's parameters start from.  The coefficients of a lattice point depend
// only on the face class (valence), so they are evaluated once per class.
#include <cmath>
#include <string>
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

// In-place inverse of a symmetric positive-definite matrix: lower Cholesky
// factor L (row-oriented, A = L L^T), then L^-1 by forward substitution,
// then A^-1 = L^-T L^-1.  Returns 0 or the failing pivot + 1.
int spd_inverse(int n, std::vector<double> &A)
{
    std::vector<double> L((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        const double *Li = &L[(size_t)i * n];
        for (int j = 0; j <= i; j++) {
            const double *Lj = &L[(size_t)j * n];
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= Li[k] * Lj[k];
            if (i == j) {
                if (!(s > 0.0)) return i + 1;
                L[(size_t)i * n + i] = std::sqrt(s);
            } else {
                L[(size_t)i * n + j] = s / Lj[j];
            }
        }
    }
    // X = L^-1, row i: X[i][j] = (delta_ij - sum_{k<i} L[i][k] X[k][j]) / L[i][i]
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        double *Xi = &X[(size_t)i * n];
        for (int k = 0; k < i; k++) {
            const double l = L[(size_t)i * n + k];
            if (l == 0.0) continue;
            const double *Xk = &X[(size_t)k * n];
            for (int j = 0; j <= k; j++) Xi[j] -= l * Xk[j];
        }
        Xi[i] += 1.0;
        const double d = L[(size_t)i * n + i];
        for (int j = 0; j <= i; j++) Xi[j] /= d;
    }
    // A^-1 = X^T X: (i, j) = sum_{k >= max(i, j)} X[k][i] X[k][j]
    std::fill(A.begin(), A.end(), 0.0);
    for (int k = 0; k < n; k++) {
        const double *Xk = &X[(size_t)k * n];
        for (int i = 0; i <= k; i++) {
            const double xi = Xk[i];
            if (xi == 0.0) continue;
            double *Ai = &A[(size_t)i * n];
            for (int j = 0; j <= i; j++) Ai[j] += xi * Xk[j];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) A[(size_t)j * n + i] = A[(size_t)i * n + j];
    return 0;
}

}  // namespace

int effm_points_per_face() { return (kPlim + 1) * (kPlim + 2) / 2; }

int effm_mass_matrix(const Topology &T, const Plan &P, double mass_per_pt, std::vector<double> &M, std::string &err)
{
    const int nv = T.nv;
    M.assign((size_t)nv * nv, 0.0);
    for (const FaceClass &C : P.classes) {
        const int nf = C.nf(), ncp = C.ncp;
        if (nf == 0) continue;
        if (C.valence != 6 && (int)C.corner.size() < ncp) {
            err = "face class without corner weights";
            return -1;
        }
        int npts = 0;
        const std::vector<double> cf = lattice_coeffs(C, npts);
        std::vector<int> ids(ncp);
        for (int f = 0; f < nf; f++) {
            for (int x = 0; x < ncp; x++) ids[x] = C.cp[(size_t)x * nf + f];
            for (int p = 0; p < npts; p++) {
                const double *c = &cf[(size_t)p * ncp];
                for (int a = 0; a < ncp; a++) {
                    const double ma = c[a] * mass_per_pt;
                    double *row = &M[(size_t)ids[a] * nv];
                    for (int b = 0; b < ncp; b++) row[ids[b]] += ma * c[b];
                }
            }
        }
    }
    return 0;
}

int effm_inverse_csr(const Topology &T, const Plan &P, double mass_per_pt, double rel_tol, EffmCsr &out,
                     std::string &err)
{
    const int nv = T.nv;
    std::vector<double> A;
    if (effm_mass_matrix(T, P, mass_per_pt, A, err)) return -1;
    const int bad = spd_inverse(nv, A);
    if (bad) {
        err = "effective mass matrix is not positive definite (pivot " + std::to_string(bad - 1) + ")";
        return -2;
    }
    out.rowptr.assign(nv + 1, 0);
    out.col.clear();
    out.val.clear();
    for (int i = 0; i < nv; i++) {
        const double di = A[(size_t)i * nv + i];
        for (int j = 0; j < nv; j++) {
            const double x = A[(size_t)i * nv + j];
            const double dj = A[(size_t)j * nv + j];
            if (std::fabs(x) / std::sqrt(std::fabs(di * dj)) > rel_tol) {
                out.col.push_back(j);
                out.val.push_back(x);
            }
        }
        out.rowptr[i + 1] = (int)out.col.size();
    }
    out.dense.swap(A);
    return 0;
}

}  // namespace umcp
