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
// The scalable form (umcp_effm_build_sparse) keeps M itself, sparse, and
// applies M^-1 by a Chebyshev solve preconditioned with D^-1 (D = diag M).
// Here: M in ELL layout for the device, the spectral interval of D^-1 M and
// the iteration's coefficients.  The interval's upper end is the Gershgorin
// bound of D^-1/2 M D^-1/2 (a true bound: above it the iteration would
// diverge); the lower end is 0.95 times the smallest Ritz value of a Lanczos
// run (Ritz values lie above the smallest eigenvalue; an interval that cuts
// off the bottom of the spectrum only slows those components; 160 steps
// resolve the bottom of a spectrum this well conditioned).  A mesh mass
// matrix is spectrally close to its diagonal: D^-1 M lies in [0.087, 3.78]
// on icospheres of 642 to 10242 vertices, a condition number near 43 that
// does not grow with the mesh.
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

std::vector<int> effm_rcm_order(const std::vector<int> &rowptr, const std::vector<int> &col)
{
    // reverse Cuthill-McKee: breadth-first from a vertex of least degree
    // (per connected component), neighbours by ascending degree then index,
    // the whole order reversed.  Rows of one wave then touch a narrow band of
    // the vectors, so the neighbours' values share cache lines.
    const int n = (int)rowptr.size() - 1;
    std::vector<int> deg((size_t)n), order;
    order.reserve((size_t)n);
    for (int i = 0; i < n; i++) deg[i] = rowptr[i + 1] - rowptr[i];
    std::vector<char> seen((size_t)n, 0);
    std::vector<int> by_deg((size_t)n);
    for (int i = 0; i < n; i++) by_deg[i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> nb;
    for (int start : by_deg) {
        if (seen[start]) continue;
        seen[start] = 1;
        size_t head = order.size();
        order.push_back(start);
        while (head < order.size()) {
            const int v = order[head++];
            nb.clear();
            for (int e = rowptr[v]; e < rowptr[v + 1]; e++)
                if (!seen[col[e]]) {
                    seen[col[e]] = 1;
                    nb.push_back(col[e]);
                }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

void effm_ell(const std::vector<int> &rowptr, const std::vector<int> &col, const std::vector<double> &val,
              const std::vector<int> &perm, int &width, std::vector<int> &ell_col, std::vector<double> &ell_val,
              std::vector<double> &dinv)
{
    const int n = (int)rowptr.size() - 1;
    std::vector<int> inv((size_t)n);
    for (int p = 0; p < n; p++) inv[perm[p]] = p;
    width = 0;
    for (int i = 0; i < n; i++) width = std::max(width, rowptr[i + 1] - rowptr[i]);
    ell_col.assign((size_t)width * n, 0);
    ell_val.assign((size_t)width * n, 0.0);
    dinv.assign((size_t)n, 0.0);
    std::vector<std::pair<int, double>> row;
    for (int p = 0; p < n; p++) {
        const int i = perm[p];
        row.clear();
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) row.push_back({inv[col[e]], val[e]});
        std::sort(row.begin(), row.end());
        int k = 0;
        for (const auto &cv : row) {
            ell_col[(size_t)k * n + p] = cv.first;
            ell_val[(size_t)k * n + p] = cv.second;
            if (cv.first == p) dinv[p] = 1.0 / cv.second;
            k++;
        }
        for (; k < width; k++) ell_col[(size_t)k * n + p] = p;   // padding: a zero on the diagonal
    }
}

namespace {

// number of eigenvalues of the symmetric tridiagonal (a, b) below x (Sturm)
int sturm_below(const std::vector<double> &a, const std::vector<double> &b, double x)
{
    int cnt = 0;
    double q = 1.0;
    for (size_t k = 0; k < a.size(); k++) {
        const double bb = k ? b[k - 1] * b[k - 1] : 0.0;
        q = a[k] - x - (k ? bb / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) cnt++;
    }
    return cnt;
}

}  // namespace

int effm_jacobi_interval(const std::vector<int> &rowptr, const std::vector<int> &col, const std::vector<double> &val,
                         double &lo, double &hi)
{
    const int n = (int)rowptr.size() - 1;
    if (n <= 0) return -1;
    std::vector<double> sd((size_t)n);
    for (int i = 0; i < n; i++) {
        double d = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (col[e] == i) d = val[e];
        if (!(d > 0.0)) return -1;
        sd[i] = 1.0 / std::sqrt(d);
    }
    hi = 0.0;
    for (int i = 0; i < n; i++) {
        double g = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) g += std::fabs(val[e]) * sd[i] * sd[col[e]];
        hi = std::max(hi, g);
    }
    // Lanczos on J = D^-1/2 M D^-1/2 from a fixed pseudo-random start
    const int steps = std::min(n, 160);
    std::vector<double> v((size_t)n), vp((size_t)n, 0.0), w((size_t)n);
    unsigned long long z = 0x9E3779B97F4A7C15ull;
    double nrm = 0.0;
    for (int i = 0; i < n; i++) {
        z ^= z << 13, z ^= z >> 7, z ^= z << 17;
        v[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        nrm += v[i] * v[i];
    }
    for (double &x : v) x /= std::sqrt(nrm);
    std::vector<double> al, be;
    double beta = 0.0;
    for (int k = 0; k < steps; k++) {
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int e = rowptr[i]; e < rowptr[i + 1]; e++) s += val[e] * sd[i] * sd[col[e]] * v[col[e]];
            w[i] = s;
        }
        double a = 0.0;
        for (int i = 0; i < n; i++) a += w[i] * v[i];
        al.push_back(a);
        double bn = 0.0;
        for (int i = 0; i < n; i++) {
            w[i] -= a * v[i] + beta * vp[i];
            bn += w[i] * w[i];
        }
        beta = std::sqrt(bn);
        if (k + 1 == steps || beta < 1e-12) break;
        be.push_back(beta);
        for (int i = 0; i < n; i++) {
            vp[i] = v[i];
            v[i] = w[i] / beta;
        }
    }
    // smallest eigenvalue of the tridiagonal by bisection on [0, hi]
    double x0 = 0.0, x1 = hi;
    for (int it = 0; it < 200 && x1 - x0 > 1e-14 * hi; it++) {
        const double xm = 0.5 * (x0 + x1);
        if (sturm_below(al, be, xm) >= 1)
            x1 = xm;
        else
            x0 = xm;
    }
    lo = 0.95 * x0;
    return lo > 0.0 && hi > lo ? 0 : -1;
}

int effm_cheb_iterations(double lo, double hi, double tol)
{
    // error bound of the Chebyshev iteration on [lo, hi]: 2 q^k in the
    // energy norm, q = (sqrt(kappa) - 1) / (sqrt(kappa) + 1); a factor
    // sqrt(kappa) more for the relative residual (and error) in the 2-norm
    const double kappa = hi / lo, sk = std::sqrt(kappa);
    const double q = (sk - 1.0) / (sk + 1.0);
    const double k = std::log(tol / (2.0 * sk)) / std::log(q);
    return std::max(1, (int)std::ceil(k));
}

void effm_cheb_coeffs(double lo, double hi, int iters, std::vector<double> &c1, std::vector<double> &c2,
                      double &inv_theta)
{
    // Chebyshev acceleration (Saad, Iterative Methods for Sparse Linear
    // Systems, 2nd ed., Algorithm 12.1, with the preconditioner D^-1):
    //   d_0 = D^-1 r_0 / theta;  x += d_k;  r -= M d_k;
    //   rho_{k+1} = 1 / (2 sigma - rho_k);
    //   d_{k+1} = rho_{k+1} rho_k d_k + (2 rho_{k+1} / delta) D^-1 r_{k+1}
    // theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma = theta / delta,
    // rho_0 = 1 / sigma
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    inv_theta = 1.0 / theta;
    c1.resize((size_t)iters);
    c2.resize((size_t)iters);
    double rho = 1.0 / sigma;
    for (int k = 0; k < iters; k++) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c1[(size_t)k] = rn * rho;
        c2[(size_t)k] = 2.0 * rn / delta;
        rho = rn;
    }
}

}  // namespace umcp
