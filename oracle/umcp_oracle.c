/*
 * umcp_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the UMCP
 * membrane-surface hot path.  See umcp_oracle.h for the contract: only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load this file's
 * library, as the checker.  The product never links it.
 *
 * It restates, in plain serial FP64 C, what the reference computes:
 *   - QUADPACK QAGS (the algorithm GSL's gsl_integration_qags implements;
 *     GSL is an un-vendored dependency of the reference, configure.ac:65-67,
 *     version unpinned) and Golub-Welsch nodes (src/gauss.C:27-299);
 *   - .mesh parsing and the combinatorial orientation of loadLattice /
 *     orientSurface (src/interpLib.C:1144-2606, 7190-7600, 8279-8328,
 *     9558-9573);
 *   - generatePlan's per-Gauss-point formula tables (src/interpLib.C:
 *     4998-6670), irregular faces by Stam's sub-patch evaluation with the
 *     subdivision matrix power A^(domain-1) taken by repeated multiplication
 *     (the reference uses LAPACK dgeev/dgetri eigen-powers, equal in exact
 *     arithmetic; LAPACK is absent here);
 *   - setg0 with the lipid packing of src/interpLib.C:2969-3388 for one
 *     type (src/lipid_composition.C:18-34, 285-302) or a composition of
 *     several (src/lipid_composition.C:47-333, drawn with a restatement of
 *     glibc's TYPE_3 rand(), which the reference calls), and the per-type
 *     energy (src/interpLib.C:3576-3610, 3805-3840);
 *   - fenergy / ifenergy summed in triangle order (src/interpLib.C:3426-3881,
 *     4180-4352; src/parallel.C:319-326);
 *   - grad / igrad through the reference's own computational graph
 *     (principal curvatures c1, c2 with the SAFETY-regularised square root,
 *     src/grad.C:65-942, 6445-7457);
 *   - the per-vertex overdamped Euler-Maruyama Brownian step
 *     r <- r - (D dt / kT) grad E + sqrt(2 D dt) xi (docs/timescales.tex:59-67,
 *     src/bd.C:1035-1036);
 *   - the global area and volume restraints (src/interpLib.C:4322-4326,
 *     src/minimize.C:90-114, src/volume.C:81-430), the hd.C effective-mass
 *     matrix, inverse, truncation and Brownian/Langevin steps
 *     (src/hd.C:418-481, 1543-1720; src/modes.C:1719-1960), evaluation at
 *     arbitrary (face, u, v) points and the particle terms (src/grad.C:
 *     3487-5690, src/interpLib.C:8330-).
 *
 * Each part is pinned by the tests against the reference's fixtures and
 * against closed forms, invariants and finite differences.  Two results are
 * pinned only to the mathematics, not bit for bit, because the libraries the
 * reference calls are absent: the irregular tables' matrix power (LAPACK's
 * eigen-route equals repeated multiplication in exact arithmetic) and QAGS
 * (GSL's output itself; the restatement is checked against exact moments).
 */
#include "umcp_oracle.h"

#include <float.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OMAXV 16          /* max valence held; the reference accepts 5..7 */
#define OMAXCP 16         /* max control points per face (val + 6) */
#define SAFETY_ (1e-12)   /* src/interp.h:14 */

/* ------------------------------------------------------------------------ */
/* QUADPACK QAGS restatement (Piessens, de Doncker-Kapenga, Ueberhuber,     */
/* Kahaner 1983: dqagse, dqk21, dqpsrt, dqelg)                               */
/* ------------------------------------------------------------------------ */

static const double XGK21[11] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
    0.000000000000000000000000000000000};
static const double WGK21[11] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077208980939000, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
    0.149445554002916905664936468389821};
static const double WG10[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};

static void qk21(double (*f)(double, void *), void *ctx, double a, double b,
                 double *result, double *abserr, double *resabs, double *resasc)
{
    const double epmach = DBL_EPSILON, uflow = DBL_MIN;
    double centr = 0.5 * (a + b), hlgth = 0.5 * (b - a), dhlgth = fabs(hlgth);
    double fv1[10], fv2[10];
    double resg = 0.0, fc = f(centr, ctx);
    double resk = WGK21[10] * fc;
    double rabs = fabs(resk);
    for (int j = 0; j < 5; j++) {
        int jtw = 2 * j + 1; /* Fortran 2j (even) -> 0-based odd */
        double absc = hlgth * XGK21[jtw];
        double f1 = f(centr - absc, ctx), f2 = f(centr + absc, ctx);
        fv1[jtw] = f1; fv2[jtw] = f2;
        double fsum = f1 + f2;
        resg += WG10[j] * fsum;
        resk += WGK21[jtw] * fsum;
        rabs += WGK21[jtw] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 5; j++) {
        int jtwm1 = 2 * j;
        double absc = hlgth * XGK21[jtwm1];
        double f1 = f(centr - absc, ctx), f2 = f(centr + absc, ctx);
        fv1[jtwm1] = f1; fv2[jtwm1] = f2;
        double fsum = f1 + f2;
        resk += WGK21[jtwm1] * fsum;
        rabs += WGK21[jtwm1] * (fabs(f1) + fabs(f2));
    }
    double reskh = resk * 0.5;
    double rasc = WGK21[10] * fabs(fc - reskh);
    for (int j = 0; j < 10; j++)
        rasc += WGK21[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
    *result = resk * hlgth;
    rabs *= dhlgth;
    rasc *= dhlgth;
    double err = fabs((resk - resg) * hlgth);
    if (rasc != 0.0 && err != 0.0) {
        double t = pow(200.0 * err / rasc, 1.5);
        err = rasc * (t < 1.0 ? t : 1.0);
    }
    if (rabs > uflow / (50.0 * epmach)) {
        double t = epmach * 50.0 * rabs;
        if (t > err) err = t;
    }
    *abserr = err;
    *resabs = rabs;
    *resasc = rasc;
}

/* keeps iord[] ordered by decreasing error; 1-based indices as in QUADPACK */
static void qpsrt(int limit, int last, int *maxerr, double *ermax, const double *elist,
                  int *iord, int *nrmax)
{
    int i, isucc, k;
    if (last <= 2) {
        iord[1] = 1;
        iord[2] = 2;
        goto done;
    }
    {
        double errmax = elist[*maxerr];
        if (*nrmax != 1) {
            int ido = *nrmax - 1;
            for (i = 1; i <= ido; i++) {
                isucc = iord[*nrmax - 1];
                if (errmax <= elist[isucc]) break;
                iord[*nrmax] = isucc;
                (*nrmax)--;
            }
        }
        int jupbn = last;
        if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
        double errmin = elist[last];
        int jbnd = jupbn - 1;
        int ibeg = *nrmax + 1;
        int placed = 0;
        if (ibeg <= jbnd) {
            for (i = ibeg; i <= jbnd; i++) {
                isucc = iord[i];
                if (errmax >= elist[isucc]) {
                    /* insert errmax by traversing the list top-down */
                    iord[i - 1] = *maxerr;
                    k = jbnd;
                    int j;
                    int put = 0;
                    for (j = i; j <= jbnd; j++) {
                        isucc = iord[k];
                        if (errmin < elist[isucc]) {
                            iord[k + 1] = last;
                            put = 1;
                            break;
                        }
                        iord[k + 1] = isucc;
                        k--;
                    }
                    if (!put) iord[i] = last;
                    placed = 1;
                    break;
                }
                iord[i - 1] = isucc;
            }
        }
        if (!placed) {
            iord[jbnd] = *maxerr;
            iord[jupbn] = last;
        }
    }
done:
    *maxerr = iord[*nrmax];
    *ermax = elist[*maxerr];
}

/* Wynn epsilon algorithm; epstab is 1-based with room for 52 entries */
static void qelg(int *n, double *epstab, double *result, double *abserr, double *res3la,
                 int *nres)
{
    const double epmach = DBL_EPSILON, oflow = DBL_MAX;
    (*nres)++;
    *abserr = oflow;
    *result = epstab[*n];
    if (*n < 3) goto out;
    {
        const int limexp = 50;
        epstab[*n + 2] = epstab[*n];
        int newelm = (*n - 1) / 2;
        epstab[*n] = oflow;
        int num = *n;
        int k1 = *n;
        for (int i = 1; i <= newelm; i++) {
            int k2 = k1 - 1, k3 = k1 - 2;
            double res = epstab[k1 + 2];
            double e0 = epstab[k3], e1 = epstab[k2], e2 = res;
            double e1abs = fabs(e1);
            double delta2 = e2 - e1, err2 = fabs(delta2);
            double tol2 = fmax(fabs(e2), e1abs) * epmach;
            double delta3 = e1 - e0, err3 = fabs(delta3);
            double tol3 = fmax(e1abs, fabs(e0)) * epmach;
            if (!(err2 > tol2 || err3 > tol3)) {
                *result = res;
                *abserr = err2 + err3;
                goto out;
            }
            double e3 = epstab[k1];
            epstab[k1] = e1;
            double delta1 = e1 - e3, err1 = fabs(delta1);
            double tol1 = fmax(e1abs, fabs(e3)) * epmach;
            double ss = 0.0;
            int omit = (err1 <= tol1 || err2 <= tol2 || err3 <= tol3);
            if (!omit) {
                ss = 1.0 / delta1 + 1.0 / delta2 - 1.0 / delta3;
                double epsinf = fabs(ss * e1);
                if (!(epsinf > 1e-4)) omit = 1;
            }
            if (omit) {
                *n = i + i - 1;
                break;
            }
            res = e1 + 1.0 / ss;
            epstab[k1] = res;
            k1 -= 2;
            double error = err2 + fabs(res - e2) + err3;
            if (error > *abserr) continue;
            *abserr = error;
            *result = res;
        }
        if (*n == limexp) *n = 2 * (limexp / 2) - 1;
        int ib = ((num / 2) * 2 == num) ? 2 : 1;
        int ie = newelm + 1;
        for (int i = 1; i <= ie; i++) {
            int ib2 = ib + 2;
            epstab[ib] = epstab[ib2];
            ib = ib2;
        }
        if (num != *n) {
            int indx = num - *n + 1;
            for (int i = 1; i <= *n; i++) {
                epstab[i] = epstab[indx];
                indx++;
            }
        }
        if (*nres < 4) {
            res3la[*nres] = *result;
            *abserr = oflow;
        } else {
            *abserr = fabs(*result - res3la[3]) + fabs(*result - res3la[2]) +
                      fabs(*result - res3la[1]);
            res3la[1] = res3la[2];
            res3la[2] = res3la[3];
            res3la[3] = *result;
        }
    }
out:
    *abserr = fmax(*abserr, 5.0 * epmach * fabs(*result));
}

double orc_qags(double (*f)(double, void *), void *ctx, double a, double b, double epsabs,
                double epsrel, int limit, double *abserr_out, int *ier_out, int *last_out)
{
    const double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
    double *alist = calloc((size_t)limit + 2, sizeof(double));
    double *blist = calloc((size_t)limit + 2, sizeof(double));
    double *rlist = calloc((size_t)limit + 2, sizeof(double));
    double *elist = calloc((size_t)limit + 2, sizeof(double));
    int *iord = calloc((size_t)limit + 2, sizeof(int));
    double rlist2[56], res3la[4];
    int ier = 0, last = 0, ierro = 0;
    double result = 0.0, abserr = 0.0, defabs, resabs;
    double area = 0.0, errsum, errbnd, errmax, small = 0.0, erlarg = 0.0, ertest = 0.0,
           correc = 0.0;
    int maxerr, nrmax, nres, numrl2, ktmin, extrap, noext, iroff1, iroff2, iroff3, ksgn;

    alist[1] = a;
    blist[1] = b;
    qk21(f, ctx, a, b, &result, &abserr, &defabs, &resabs);
    double dres = fabs(result);
    errbnd = fmax(epsabs, epsrel * dres);
    last = 1;
    rlist[1] = result;
    elist[1] = abserr;
    iord[1] = 1;
    if (abserr <= 100.0 * epmach * defabs && abserr > errbnd) ier = 2;
    if (limit == 1) ier = 1;
    if (ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0.0) goto finish;

    rlist2[1] = result;
    errmax = abserr;
    maxerr = 1;
    area = result;
    errsum = abserr;
    abserr = oflow;
    nrmax = 1;
    nres = 0;
    numrl2 = 2;
    ktmin = 0;
    extrap = 0;
    noext = 0;
    iroff1 = iroff2 = iroff3 = 0;
    ksgn = -1;
    if (dres >= (1.0 - 50.0 * epmach) * defabs) ksgn = 1;

    for (last = 2; last <= limit; last++) {
        double a1 = alist[maxerr];
        double b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
        double a2 = b1, b2 = blist[maxerr];
        double erlast = errmax;
        double area1, error1, defab1, area2, error2, defab2, dum;
        qk21(f, ctx, a1, b1, &area1, &error1, &dum, &defab1);
        qk21(f, ctx, a2, b2, &area2, &error2, &dum, &defab2);
        double area12 = area1 + area2, erro12 = error1 + error2;
        errsum = errsum + erro12 - errmax;
        area = area + area12 - rlist[maxerr];
        if (!(defab1 == error1 || defab2 == error2)) {
            if (!(fabs(rlist[maxerr] - area12) > 1e-5 * fabs(area12) || erro12 < 0.99 * errmax)) {
                if (extrap) iroff2++;
                else iroff1++;
            }
            if (last > 10 && erro12 > errmax) iroff3++;
        }
        rlist[maxerr] = area1;
        rlist[last] = area2;
        errbnd = fmax(epsabs, epsrel * fabs(area));
        if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
        if (iroff2 >= 5) ierro = 3;
        if (last == limit) ier = 1;
        if (fmax(fabs(a1), fabs(b2)) <= (1.0 + 100.0 * epmach) * (fabs(a2) + 1000.0 * uflow))
            ier = 4;
        if (error2 > error1) {
            alist[maxerr] = a2;
            alist[last] = a1;
            blist[last] = b1;
            rlist[maxerr] = area2;
            rlist[last] = area1;
            elist[maxerr] = error2;
            elist[last] = error1;
        } else {
            alist[last] = a2;
            blist[maxerr] = b1;
            blist[last] = b2;
            elist[maxerr] = error1;
            elist[last] = error2;
        }
        qpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
        if (errsum <= errbnd) goto sum_up;
        if (ier != 0) goto tail;
        if (last == 2) {
            small = fabs(b - a) * 0.375;
            erlarg = errsum;
            ertest = errbnd;
            rlist2[2] = area;
            continue;
        }
        if (noext) continue;
        erlarg -= erlast;
        if (fabs(b1 - a1) > small) erlarg += erro12;
        if (!extrap) {
            if (fabs(blist[maxerr] - alist[maxerr]) > small) continue;
            extrap = 1;
            nrmax = 2;
        }
        if (ierro != 3 && erlarg > ertest) {
            int id = nrmax, jupbnd = last, skip = 0;
            if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
            for (int k = id; k <= jupbnd; k++) {
                maxerr = iord[nrmax];
                errmax = elist[maxerr];
                if (fabs(blist[maxerr] - alist[maxerr]) > small) {
                    skip = 1;
                    break;
                }
                nrmax++;
            }
            if (skip) continue;
        }
        numrl2++;
        rlist2[numrl2] = area;
        {
            double reseps, abseps;
            qelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
            ktmin++;
            if (ktmin > 5 && abserr < 1e-3 * errsum) ier = 5;
            if (abseps < abserr) {
                ktmin = 0;
                abserr = abseps;
                result = reseps;
                correc = erlarg;
                ertest = fmax(epsabs, epsrel * fabs(reseps));
                if (abserr <= ertest) goto tail;
            }
        }
        if (numrl2 == 1) noext = 1;
        if (ier == 5) goto tail;
        maxerr = iord[1];
        errmax = elist[maxerr];
        nrmax = 1;
        extrap = 0;
        small *= 0.5;
        erlarg = errsum;
    }
    if (last > limit) last = limit;

tail:
    if (abserr == oflow) goto sum_up;
    if (ier + ierro != 0) {
        if (ierro == 3) abserr += correc;
        if (ier == 0) ier = 3;
        if (result != 0.0 && area != 0.0) {
            if (abserr / fabs(result) > errsum / fabs(area)) goto sum_up;
        } else if (abserr > errsum) {
            goto sum_up;
        } else if (area == 0.0) {
            goto final;
        }
    }
    if (ksgn == -1 && fmax(fabs(result), fabs(area)) <= defabs * 0.01) goto final;
    if (0.01 > (result / area) || (result / area) > 100.0 || errsum > fabs(area)) ier = 6;
    goto final;
sum_up:
    result = 0.0;
    for (int k = 1; k <= last; k++) result += rlist[k];
    abserr = errsum;
final:
    if (ier > 2) ier--;
finish:
    if (abserr_out) *abserr_out = abserr;
    if (ier_out) *ier_out = ier;
    if (last_out) *last_out = last;
    free(alist); free(blist); free(rlist); free(elist); free(iord);
    return result;
}

/* ------------------------------------------------------------------------ */
/* Golub-Welsch nodes for the two singular weights of src/gauss.C          */
/* ------------------------------------------------------------------------ */

typedef struct {
    double k, v;
    int n, type;
    const double *a, *b;
} gw_block;

static double monic_p(const gw_block *B, double x)
{
    double pnm1 = 0.0, pn = 1.0;
    for (int i = 0; i < B->n; i++) {
        double p = (x - B->a[i]) * pn - B->b[i] * pnm1;
        pnm1 = pn;
        pn = p;
    }
    return pn;
}

/* weight in u at fixed v: ((u+v) k^{log2(u+v)})^4 (src/gauss.C:27-59) */
static double gw_fn_u(double u, void *ctx)
{
    const gw_block *B = ctx;
    double val = (1.0 / (u + B->v)) * pow(B->k, -log2(u + B->v));
    double pn = monic_p(B, u);
    if (B->type == 1) return u * pn * pn / (val * val * val * val);
    return pn * pn / (val * val * val * val);
}

/* weight in v: (1 - k^{4 log2 v} v^5) ln2 / (ln32 + 4 ln k) (src/gauss.C:61-94) */
static double gw_fn_v(double v, void *ctx)
{
    const gw_block *B = ctx;
    double k = B->k;
    double val = -(-1 + pow(k, 4.0 * log2(v)) * v * v * v * v * v) * log(2.0) /
                 (log(32.0) + 4.0 * log(k));
    double pn = monic_p(B, v);
    if (B->type == 1) return v * pn * pn * val;
    return pn * pn * val;
}

/* symmetric eigen-decomposition by cyclic Jacobi rotations (n <= 8); on
 * return d[] ascending, V columns (V[r*n+c]) are the unit eigenvectors */
static void jacobi_eig(int n, double *A, double *d, double *V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) d[i] = A[i * n + i];
    for (int i = 0; i < n; i++) /* selection sort ascending, permuting columns */
        for (int j = i + 1; j < n; j++)
            if (d[j] < d[i]) {
                double t = d[i]; d[i] = d[j]; d[j] = t;
                for (int k = 0; k < n; k++) {
                    double tv = V[k * n + i]; V[k * n + i] = V[k * n + j]; V[k * n + j] = tv;
                }
            }
}

/* recurrence by integrating monic polynomials with QAGS(epsabs 1e-9, epsrel 0,
 * limit 100), then Golub-Welsch (src/gauss.C:96-197 / 199-299) */
static void golub_welsch(double (*fn)(double, void *), double k, double v, double hi,
                         int npts, double *x, double *w)
{
    double a[16] = {0}, b[16] = {0};
    gw_block B = {k, v, 0, 0, a, b};
    double prev = 1.0, mu0 = 0.0;
    for (int i = 0; i <= npts; i++) {
        B.n = i;
        B.type = 1;
        double num = orc_qags(fn, &B, 0.0, hi, 1e-9, 0.0, 100, NULL, NULL, NULL);
        B.type = 0;
        double den = orc_qags(fn, &B, 0.0, hi, 1e-9, 0.0, 100, NULL, NULL, NULL);
        a[i] = num / den;
        b[i] = (i > 0) ? den / prev : 0.0;
        if (i == 0) mu0 = den;
        prev = den;
    }
    double J[64] = {0}, V[64], d[8];
    for (int i = 0; i < npts; i++) {
        J[i * npts + i] = a[i];
        if (i > 0) J[(i - 1) * npts + i] = J[i * npts + i - 1] = sqrt(b[i]);
    }
    jacobi_eig(npts, J, d, V);
    for (int i = 0; i < npts; i++) {
        double nrm = 0.0;
        for (int j = 0; j < npts; j++) nrm += V[j * npts + i] * V[j * npts + i];
        x[i] = d[i];
        w[i] = mu0 * V[0 * npts + i] * V[0 * npts + i] / nrm;
    }
}

void orc_get_v_pts(double *v_pts, double *w_pts, int npts, double k)
{
    golub_welsch(gw_fn_v, k, 0.0, 1.0, npts, v_pts, w_pts);
}

void orc_get_u_pts(double *u_pts, double *w_pts, int npts, double k, double v)
{
    golub_welsch(gw_fn_u, k, v, 1.0 - v, npts, u_pts, w_pts);
}

/* ------------------------------------------------------------------------ */
/* Quartic box-spline basis (Loop), src/interpLib.C:5836-5918             */
/* ------------------------------------------------------------------------ */

/* each basis n_i = (1/12) sum c * u^p v^q w^r with w = 1-u-v; terms listed
 * as {c,p,q,r}; 0 terminates */
static const int BOX_TERMS[12][16][4] = {
    /* n1 */ {{1, 4, 0, 0}, {2, 3, 1, 0}},
    /* n2 */ {{1, 4, 0, 0}, {2, 3, 0, 1}},
    /* n3 */ {{1, 4, 0, 0}, {2, 3, 0, 1}, {6, 3, 1, 0}, {6, 2, 1, 1}, {12, 2, 2, 0},
              {6, 1, 2, 1}, {6, 1, 3, 0}, {2, 0, 3, 1}, {1, 0, 4, 0}},
    /* n4 */ {{6, 4, 0, 0}, {24, 3, 0, 1}, {24, 2, 0, 2}, {8, 1, 0, 3}, {1, 0, 0, 4},
              {24, 3, 1, 0}, {60, 2, 1, 1}, {36, 1, 1, 2}, {6, 0, 1, 3}, {24, 2, 2, 0},
              {36, 1, 2, 1}, {12, 0, 2, 2}, {8, 1, 3, 0}, {6, 0, 3, 1}, {1, 0, 4, 0}},
    /* n5 */ {{1, 4, 0, 0}, {6, 3, 0, 1}, {12, 2, 0, 2}, {6, 1, 0, 3}, {1, 0, 0, 4},
              {2, 3, 1, 0}, {6, 2, 1, 1}, {6, 1, 1, 2}, {2, 0, 1, 3}},
    /* n6 */ {{2, 1, 3, 0}, {1, 0, 4, 0}},
    /* n7 */ {{1, 4, 0, 0}, {6, 3, 0, 1}, {12, 2, 0, 2}, {6, 1, 0, 3}, {1, 0, 0, 4},
              {8, 3, 1, 0}, {36, 2, 1, 1}, {36, 1, 1, 2}, {8, 0, 1, 3}, {24, 2, 2, 0},
              {60, 1, 2, 1}, {24, 0, 2, 2}, {24, 1, 3, 0}, {24, 0, 3, 1}, {6, 0, 4, 0}},
    /* n8 */ {{1, 4, 0, 0}, {8, 3, 0, 1}, {24, 2, 0, 2}, {24, 1, 0, 3}, {6, 0, 0, 4},
              {6, 3, 1, 0}, {36, 2, 1, 1}, {60, 1, 1, 2}, {24, 0, 1, 3}, {12, 2, 2, 0},
              {36, 1, 2, 1}, {24, 0, 2, 2}, {6, 1, 3, 0}, {8, 0, 3, 1}, {1, 0, 4, 0}},
    /* n9 */ {{2, 1, 0, 3}, {1, 0, 0, 4}},
    /* n10 */ {{2, 0, 3, 1}, {1, 0, 4, 0}},
    /* n11 */ {{2, 1, 0, 3}, {1, 0, 0, 4}, {6, 1, 1, 2}, {6, 0, 1, 3}, {6, 1, 2, 1},
               {12, 0, 2, 2}, {2, 1, 3, 0}, {6, 0, 3, 1}, {1, 0, 4, 0}},
    /* n12 */ {{1, 0, 0, 4}, {2, 0, 1, 3}},
};
/* point_list slot y carries basis CEFF[y] (the ceff_map permutation) */
static const int CEFF[12] = {8, 7, 4, 5, 9, 12, 11, 10, 6, 3, 1, 2};

static double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* monomial coefficients m[a][b] of u^a v^b (a+b <= 4), times 12 */
static void box_monomials(int which, double m[5][5])
{
    memset(m, 0, sizeof(double) * 25);
    for (int t = 0; t < 16 && BOX_TERMS[which][t][0]; t++) {
        int c = BOX_TERMS[which][t][0], p = BOX_TERMS[which][t][1];
        int q = BOX_TERMS[which][t][2], r = BOX_TERMS[which][t][3];
        /* w^r = sum_{i+j+l=r} r!/(i!j!l!) (-u)^i (-v)^j */
        for (int i = 0; i <= r; i++)
            for (int j = 0; i + j <= r; j++) {
                double mult = binom(r, i) * binom(r - i, j) * (((i + j) & 1) ? -1.0 : 1.0);
                m[p + i][q + j] += c * mult;
            }
    }
}

/* out[d][y]: d = 0 value, 1 u, 2 v, 3 uu, 4 uv, 5 vv, 6 uuu, 7 uuv, 8 uvv,
 * 9 vvv at (u,v) in slot order (the third derivatives are what pgrad's
 * (u, v) gradient needs, src/grad.C:3620-3700) */
void orc_basis10(double u, double v, double out[10][12])
{
    double up[5] = {1, u, u * u, u * u * u, u * u * u * u};
    double vp[5] = {1, v, v * v, v * v * v, v * v * v * v};
    /* derivative orders (du, dv) of the ten rows */
    static const int DU[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0};
    static const int DV[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
    for (int y = 0; y < 12; y++) {
        double m[5][5];
        box_monomials(CEFF[y] - 1, m);
        double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a <= 4; a++)
            for (int b = 0; a + b <= 4; b++) {
                double c = m[a][b] / 12.0;
                if (c == 0.0) continue;
                for (int d = 0; d < 10; d++) {
                    int i = DU[d], j = DV[d];
                    if (a < i || b < j) continue;
                    double fa = 1.0, fb = 1.0;
                    for (int t = 0; t < i; t++) fa *= (a - t);
                    for (int t = 0; t < j; t++) fb *= (b - t);
                    s[d] += c * fa * fb * up[a - i] * vp[b - j];
                }
            }
        for (int d = 0; d < 10; d++) out[d][y] = s[d];
    }
}

/* out[d][y]: d = 0 value, 1 u, 2 v, 3 uu, 4 uv, 5 vv at (u,v) in slot order */
void orc_basis(double u, double v, double out[6][12])
{
    double b[10][12];
    orc_basis10(u, v, b);
    for (int d = 0; d < 6; d++)
        for (int y = 0; y < 12; y++) out[d][y] = b[d][y];
}

/* ------------------------------------------------------------------------ */
/* Stam evaluation near an extraordinary vertex, src/interpLib.C:5011-5161, */
/* 5687-5809                                                                */
/* ------------------------------------------------------------------------ */

/* Atot: (N+6) x N Loop subdivision stencil, N = val + 6, in point_list order */
static void subdivision_matrix(int val, double *A)
{
    int N = val + 6;
    double w = 3.0 / (8 * val), w6 = 3.0 / (8 * 6);
    memset(A, 0, sizeof(double) * (size_t)(N + 6) * N);
#define AS(r, c) A[(r) * N + (c)]
    AS(0, 0) = 1 - val * w;
    for (int p = 1; p < 1 + val; p++) AS(0, p) = w;
    AS(1, 0) = 3.0 / 8; AS(1, 1) = 3.0 / 8; AS(1, 2) = 1.0 / 8; AS(1, val) = 1.0 / 8;
    for (int p = 2; p < val; p++) {
        AS(p, 0) = 3.0 / 8; AS(p, p) = 3.0 / 8; AS(p, p - 1) = 1.0 / 8; AS(p, p + 1) = 1.0 / 8;
    }
    AS(val, 0) = 3.0 / 8; AS(val, val) = 3.0 / 8; AS(val, val - 1) = 1.0 / 8; AS(val, 1) = 1.0 / 8;
    AS(val + 1, 1) = 3.0 / 8; AS(val + 1, val) = 3.0 / 8; AS(val + 1, 0) = 1.0 / 8;
    AS(val + 1, val + 1) = 1.0 / 8;
    AS(val + 2, 1) = 1 - 6 * w6;
    AS(val + 2, 0) = w6; AS(val + 2, val) = w6; AS(val + 2, val + 1) = w6;
    AS(val + 2, val + 2) = w6; AS(val + 2, val + 3) = w6; AS(val + 2, 2) = w6;
    AS(val + 3, 1) = 3.0 / 8; AS(val + 3, 2) = 3.0 / 8; AS(val + 3, 0) = 1.0 / 8;
    AS(val + 3, val + 3) = 1.0 / 8;
    AS(val + 4, 2) = 1 - 6 * w6;
    AS(val + 4, 0) = w6; AS(val + 4, 1) = w6; AS(val + 4, val + 3) = w6;
    AS(val + 4, val + 4) = w6; AS(val + 4, val + 5) = w6; AS(val + 4, 3) = w6;
    AS(val + 5, 2) = 3.0 / 8; AS(val + 5, 3) = 3.0 / 8; AS(val + 5, 0) = 1.0 / 8;
    AS(val + 5, val + 5) = 1.0 / 8;
    AS(val + 6, 1) = 3.0 / 8; AS(val + 6, val + 1) = 3.0 / 8; AS(val + 6, val + 2) = 1.0 / 8;
    AS(val + 6, val) = 1.0 / 8;
    AS(val + 7, 1) = 3.0 / 8; AS(val + 7, val + 2) = 3.0 / 8; AS(val + 7, val + 3) = 1.0 / 8;
    AS(val + 7, val + 1) = 1.0 / 8;
    AS(val + 8, 1) = 3.0 / 8; AS(val + 8, val + 3) = 3.0 / 8; AS(val + 8, 2) = 1.0 / 8;
    AS(val + 8, val + 2) = 1.0 / 8;
    AS(val + 9, 2) = 3.0 / 8; AS(val + 9, val + 3) = 3.0 / 8; AS(val + 9, 1) = 1.0 / 8;
    AS(val + 9, val + 4) = 1.0 / 8;
    AS(val + 10, 2) = 3.0 / 8; AS(val + 10, val + 4) = 3.0 / 8; AS(val + 10, val + 5) = 1.0 / 8;
    AS(val + 10, val + 3) = 1.0 / 8;
    AS(val + 11, 2) = 3.0 / 8; AS(val + 11, val + 5) = 3.0 / 8; AS(val + 11, 3) = 1.0 / 8;
    AS(val + 11, val + 4) = 1.0 / 8;
#undef AS
}

/* out is [6][val+6]: derivative tables r_w, r_u, r_v, r_uu, r_uv, r_vv at the
 * original parameters (u, v) of an irregular face (src/interpLib.C:5687-5918) */
static int irregular_table_nd(int val, double fu, double fv, int nd, double *out)
{
    int N = val + 6;
    if (val < 3 || N > OMAXCP) return -1;
    double *Atot = malloc(sizeof(double) * (size_t)(N + 6) * N);
    double *P = malloc(sizeof(double) * (size_t)N * N);
    double *T = malloc(sizeof(double) * (size_t)N * N);
    double *Ause = malloc(sizeof(double) * (size_t)(N + 6) * N);
    subdivision_matrix(val, Atot);

    double scale_u = 1.0, scale_v = 1.0;
    double f_domain = -log10(fu + fv) / log10(2.0);
    if (fu + fv >= 1.0 - 1e-9) f_domain = 1;
    int domain = (int)lround(ceil(f_domain));
    double pow2 = pow(2.0, domain - 1.0);
    scale_u *= pow2; scale_v *= pow2;
    fu *= pow2; fv *= pow2;
    int pc[12];
    if (fu > 0.5) {
        fv = 2 * fv; fu = 2.0 * fu - 1.0;
        scale_u *= 2; scale_v *= 2;
        int t[12] = {2, val + 3, val + 4, val + 5, 3, 0, 1, val + 2, val + 8, val + 9, val + 10, val + 11};
        memcpy(pc, t, sizeof t);
    } else if (fv > 0.5) {
        fu = 2 * fu; fv = 2.0 * fv - 1.0;
        scale_u *= 2; scale_v *= 2;
        int t[12] = {1, val + 2, val + 3, 2, 0, val, val + 1, val + 6, val + 7, val + 8, val + 9, val + 4};
        memcpy(pc, t, sizeof t);
    } else {
        fv = 1 - 2 * fv; fu = 1 - 2 * fu;
        scale_u *= -2; scale_v *= -2;
        int t[12] = {val + 3, 2, 1, val + 2, val + 8, val + 9, val + 4, val + 5, 3, 0, val, val + 1};
        memcpy(pc, t, sizeof t);
    }
    /* P = A^(domain-1), A = top N x N block of Atot */
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) P[i * N + j] = (i == j);
    for (int s = 0; s < domain - 1; s++) {
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                double acc = 0.0;
                for (int l = 0; l < N; l++) acc += Atot[i * N + l] * P[l * N + j];
                T[i * N + j] = acc;
            }
        memcpy(P, T, sizeof(double) * (size_t)N * N);
    }
    for (int f = 0; f < N + 6; f++)
        for (int ii = 0; ii < N; ii++) {
            double acc = 0.0;
            for (int e = 0; e < N; e++) acc += Atot[f * N + e] * P[e * N + ii];
            Ause[f * N + ii] = acc;
        }
    double B[10][12];
    orc_basis10(fu, fv, B);
    double sc[10] = {1.0, scale_u, scale_v, scale_u * scale_u, scale_u * scale_v, scale_v * scale_v,
                     scale_u * scale_u * scale_u, scale_u * scale_u * scale_v, scale_u * scale_v * scale_v,
                     scale_v * scale_v * scale_v};
    for (int x = 0; x < N; x++)
        for (int d = 0; d < nd; d++) {
            double acc = 0.0;
            for (int y = 0; y < 12; y++) acc += Ause[pc[y] * N + x] * B[d][y] * sc[d];
            out[d * N + x] = acc;
        }
    free(Atot); free(P); free(T); free(Ause);
    return 0;
}

int orc_irregular_table(int val, double fu, double fv, double *out)
{
    return irregular_table_nd(val, fu, fv, 6, out);
}

/* ------------------------------------------------------------------------ */
/* Mesh                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int ncp;
    int cp[OMAXCP];
    double pbc[3 * OMAXCP];
    double t[6][OMAXCP];   /* r_w, r_u, r_v, r_uu, r_uv, r_vv */
    double weight, g0, RuRv0;
    int tri;
} oform;

struct omesh {
    int nv;
    double PBC[3][3];
    int fix_sense;
    double *r;        /* positions read from the file, 3 nv */
    int *val;
    int *nb;          /* nv * OMAXV: neighbour lists, finally oriented */
    int *rev;         /* edge_rev */
    /* triangles (final order) */
    int nt;
    int *tri;         /* 3 nt, ascending ids */
    int *tri_face;    /* face index of triangle t (theTriangles[t].f) */
    /* lipid composition (src/lipid_composition.C:47-333): nlt types with
     * spontaneous curvature lt_c0, area per lipid lt_apl and input parts per
     * leaflet; Lo/Li the area of each type in each triangle's outer/inner
     * leaflet, [nt][nlt]; with no composition set, one type of
     * (prm.lipid_c0, prm.lipid_apl) */
    int nlt;
    double *lt_c0, *lt_apl, *lt_in, *lt_out;
    unsigned int lt_seed;
    double *Lo, *Li;
    /* faces */
    int planned;
    int nfr, nfi;
    int *face_owner, *face_edge;
    oform *F;         /* regular: nfr*15 ; irregular after: nfi*16 */
    double *face_c0;
    oparams prm;
    /* global restraints: area (GLOBAL_AREA form) and enclosed volume */
    double KA_global, KV, A0_target, V0_target;
    double A0_setg0, V_setg0;   /* sum(dudv g0 w) and the volume at setg0 */
};

static int omesh_adjacent(const omesh *m, int a, int b)
{
    for (int e = 0; e < m->val[a]; e++)
        if (m->nb[a * OMAXV + e] == b) return 1;
    return 0;
}

/* constructTriangles enumeration (src/interpLib.C:7190-7330) */
static int build_triangles(const omesh *m, int **out)
{
    int cap = 2 * m->nv + 16, n = 0;
    int *t = malloc(sizeof(int) * 3 * (size_t)cap);
    for (int i = 0; i < m->nv; i++)
        for (int vi = 0; vi < m->val[i]; vi++) {
            int j = m->nb[i * OMAXV + vi];
            if (j < i) continue;
            for (int vj = 0; vj < m->val[j]; vj++) {
                int k = m->nb[j * OMAXV + vj];
                if (k == i || k < j) continue;
                if (!omesh_adjacent(m, k, i)) continue;
                if (n == cap) {
                    cap *= 2;
                    t = realloc(t, sizeof(int) * 3 * (size_t)cap);
                }
                t[3 * n] = i; t[3 * n + 1] = j; t[3 * n + 2] = k;
                n++;
            }
        }
    *out = t;
    return n;
}

/* incident triangles per vertex */
static void vertex_tris(const omesh *m, const int *tri, int nt, int **start, int **list)
{
    int *cnt = calloc((size_t)m->nv + 1, sizeof(int));
    for (int t = 0; t < nt; t++)
        for (int c = 0; c < 3; c++) cnt[tri[3 * t + c] + 1]++;
    for (int v = 0; v < m->nv; v++) cnt[v + 1] += cnt[v];
    int *lst = malloc(sizeof(int) * 3 * (size_t)nt);
    int *pos = malloc(sizeof(int) * (size_t)m->nv);
    memcpy(pos, cnt, sizeof(int) * (size_t)m->nv);
    for (int t = 0; t < nt; t++)
        for (int c = 0; c < 3; c++) lst[pos[tri[3 * t + c]]++] = t;
    free(pos);
    *start = cnt;
    *list = lst;
}

/* orient all triangles consistently with triangle 0 carrying sense s0, then
 * sort each vertex's neighbours so that (i, nb[e+1], nb[e]) is an oriented
 * triangle, keeping the current first neighbour first (src/interpLib.C:
 * 1781-2020, 7515-7600; the sense rule of compare_sense, 113-123) */
static int orient_and_sort(omesh *m, int s0)
{
    int *tri, nt = build_triangles(m, &tri);
    if (nt == 0) { free(tri); return -1; }
    int *vs, *vl;
    vertex_tris(m, tri, nt, &vs, &vl);
    int *oriented = malloc(sizeof(int) * 3 * (size_t)nt); /* cyclic oriented order */
    char *done = calloc((size_t)nt, 1);
    int *queue = malloc(sizeof(int) * (size_t)nt);
    int qh = 0, qt = 0, rc = 0;
    oriented[0] = tri[0];
    if (s0 >= 0) { oriented[1] = tri[1]; oriented[2] = tri[2]; }
    else { oriented[1] = tri[2]; oriented[2] = tri[1]; }
    done[0] = 1;
    queue[qt++] = 0;
    while (qh < qt) {
        int t = queue[qh++];
        for (int c = 0; c < 3; c++) {
            int p = oriented[3 * t + c], q = oriented[3 * t + (c + 1) % 3];
            /* neighbour triangle sharing edge p-q must traverse q->p */
            for (int x = vs[p]; x < vs[p + 1]; x++) {
                int t2 = vl[x];
                if (t2 == t) continue;
                int has_q = (tri[3 * t2] == q || tri[3 * t2 + 1] == q || tri[3 * t2 + 2] == q);
                if (!has_q) continue;
                int r3 = tri[3 * t2] + tri[3 * t2 + 1] + tri[3 * t2 + 2] - p - q;
                if (!done[t2]) {
                    oriented[3 * t2] = q; oriented[3 * t2 + 1] = p; oriented[3 * t2 + 2] = r3;
                    done[t2] = 1;
                    queue[qt++] = t2;
                } else {
                    int ok = 0;
                    for (int cc = 0; cc < 3; cc++)
                        if (oriented[3 * t2 + cc] == q && oriented[3 * t2 + (cc + 1) % 3] == p) ok = 1;
                    if (!ok) rc = -2; /* not orientable */
                }
            }
        }
    }
    for (int t = 0; t < nt; t++) if (!done[t]) rc = -3;
    if (rc == 0) {
        int tmp[OMAXV];
        for (int i = 0; i < m->nv && rc == 0; i++) {
            int val = m->val[i];
            int last = m->nb[i * OMAXV + 0];
            tmp[0] = last;
            for (int e = 1; e < val; e++) {
                int next = -1;
                for (int x = vs[i]; x < vs[i + 1]; x++) {
                    int t = vl[x];
                    for (int c = 0; c < 3; c++)
                        if (oriented[3 * t + c] == last && oriented[3 * t + (c + 1) % 3] == i)
                            next = oriented[3 * t + (c + 2) % 3];
                    if (next >= 0) break;
                }
                if (next < 0) { rc = -4; break; }
                tmp[e] = next;
                last = next;
            }
            if (rc == 0) memcpy(m->nb + i * OMAXV, tmp, sizeof(int) * (size_t)val);
        }
    }
    free(tri); free(vs); free(vl); free(oriented); free(done); free(queue);
    return rc;
}

static void set_edge_rev(omesh *m)
{
    for (int i = 0; i < m->nv; i++)
        for (int e = 0; e < m->val[i]; e++) {
            int j = m->nb[i * OMAXV + e];
            for (int e2 = 0; e2 < m->val[j]; e2++)
                if (m->nb[j * OMAXV + e2] == i) m->rev[i * OMAXV + e] = e2;
        }
}

static int read_line(FILE *fp, char *buf, int n)
{
    if (!fgets(buf, n, fp)) { buf[0] = 0; return 0; }
    return 1;
}

int orc_load(const char *path, omesh **out)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char buf[4096];
    omesh *m = calloc(1, sizeof(omesh));
    m->PBC[0][0] = m->PBC[1][1] = m->PBC[2][2] = 1e10;
    read_line(fp, buf, sizeof buf);
    int vec3 = (strncmp(buf, "3D", 2) == 0 || strncmp(buf, "3d", 2) == 0);
    read_line(fp, buf, sizeof buf);
    sscanf(buf, "%lf %lf %lf", &m->PBC[0][0], &m->PBC[0][1], &m->PBC[0][2]);
    read_line(fp, buf, sizeof buf);
    sscanf(buf, "%lf %lf %lf", &m->PBC[1][0], &m->PBC[1][1], &m->PBC[1][2]);
    if (vec3) {
        read_line(fp, buf, sizeof buf);
        sscanf(buf, "%lf %lf %lf", &m->PBC[2][0], &m->PBC[2][1], &m->PBC[2][2]);
    } else {
        m->PBC[2][0] = m->PBC[2][1] = 0.0;
        m->PBC[2][2] = 1e6;
    }
    m->fix_sense = 1;
    long pos = ftell(fp);
    read_line(fp, buf, sizeof buf);
    if (strncmp(buf, "sense", 5) == 0) sscanf(buf, "sense %d", &m->fix_sense);
    else fseek(fp, pos, SEEK_SET);

    int cap = 1024;
    m->r = malloc(sizeof(double) * 3 * (size_t)cap);
    m->val = malloc(sizeof(int) * (size_t)cap);
    m->nb = malloc(sizeof(int) * OMAXV * (size_t)cap);
    int rc = 0;
    while (read_line(fp, buf, sizeof buf)) {
        if (strncmp(buf, "ntri", 4) == 0) break;
        int id, val, off = 0;
        double x, y, z;
        char *s = buf;
        if (*s == 'P' || *s == 'Q') { rc = -5; break; } /* protein points: not on this path */
        if (sscanf(s, "%d %lf %lf %lf %d%n", &id, &x, &y, &z, &val, &off) < 5) continue;
        if (val < 5 || val > 7) { rc = -6; break; }   /* src/interpLib.C:1287-1291 */
        if (m->nv == cap) {
            cap *= 2;
            m->r = realloc(m->r, sizeof(double) * 3 * (size_t)cap);
            m->val = realloc(m->val, sizeof(int) * (size_t)cap);
            m->nb = realloc(m->nb, sizeof(int) * OMAXV * (size_t)cap);
        }
        m->r[3 * m->nv] = x; m->r[3 * m->nv + 1] = y; m->r[3 * m->nv + 2] = z;
        m->val[m->nv] = val;
        s += off;
        for (int e = 0; e < val; e++) {
            int n2 = 0, nbv;
            if (sscanf(s, "%d%n", &nbv, &n2) < 1) { rc = -7; break; }
            m->nb[m->nv * OMAXV + e] = nbv;
            s += n2;
        }
        if (rc) break;
        m->nv++;
    }
    fclose(fp);
    if (rc == 0 && m->nv == 0) rc = -8;
    for (int i = 0; rc == 0 && i < m->nv; i++)
        for (int e = 0; e < m->val[i]; e++) {
            int j = m->nb[i * OMAXV + e];
            if (j < 0 || j >= m->nv || j == i || !omesh_adjacent(m, j, i)) rc = -9;
        }
    if (rc) { orc_free(m); return rc; }
    m->rev = malloc(sizeof(int) * OMAXV * (size_t)m->nv);
    /* loadLattice orientation pass: triangle 0 from the file lists, sense fix_sense */
    rc = orient_and_sort(m, m->fix_sense);
    /* final triangle order: constructTriangles over the lists orientSurface starts from */
    if (rc == 0) {
        m->nt = build_triangles(m, &m->tri);
        /* orientSurface pass: triangle 0 rebuilt from the sorted lists, sense +1 */
        rc = orient_and_sort(m, 1);
    }
    if (rc) { orc_free(m); return rc - 20; }
    set_edge_rev(m);
    m->tri_face = malloc(sizeof(int) * (size_t)m->nt);
    for (int t = 0; t < m->nt; t++) m->tri_face[t] = -1;
    m->nlt = 0;
    m->lt_c0 = m->lt_apl = m->lt_in = m->lt_out = NULL;
    m->lt_seed = 1;
    m->Lo = calloc((size_t)m->nt, sizeof(double));
    m->Li = calloc((size_t)m->nt, sizeof(double));
    m->prm.kc = 14.0; m->prm.kg = 0.0; m->prm.KA = 0.0; m->prm.k_reg = 0.0;
    m->prm.mode = 0; m->prm.lipid_c0 = 0.0; m->prm.lipid_apl = 68.3;
    *out = m;
    return 0;
}

void orc_free(omesh *m)
{
    if (!m) return;
    free(m->r); free(m->val); free(m->nb); free(m->rev); free(m->tri); free(m->tri_face);
    free(m->Lo); free(m->Li); free(m->face_owner); free(m->face_edge); free(m->F);
    free(m->lt_c0); free(m->lt_apl); free(m->lt_in); free(m->lt_out);
    free(m->face_c0);
    free(m);
}

int orc_nv(const omesh *m) { return m->nv; }
int orc_ntri(const omesh *m) { return m->nt; }
int orc_tri_order_face(const omesh *m, int t) { return m->tri_face[t]; }
void orc_neighbors(const omesh *m, int v, int *val, int *nb)
{
    *val = m->val[v];
    memcpy(nb, m->nb + v * OMAXV, sizeof(int) * (size_t)m->val[v]);
}
int orc_nfaces(const omesh *m, int *nreg, int *nirr)
{
    if (nreg) *nreg = m->nfr;
    if (nirr) *nirr = m->nfi;
    return m->nfr + m->nfi;
}
void orc_get_positions(const omesh *m, double *r)
{
    memcpy(r, m->r, sizeof(double) * 3 * (size_t)m->nv);
    r[3 * m->nv] = r[3 * m->nv + 1] = r[3 * m->nv + 2] = 1.0;
}

/* src/mutil.C:541-599 */
static void min_image(double *dr, double PBC[3][3], double *put)
{
    int done = 0;
    put[0] = put[1] = put[2] = 0.0;
    double tr[3] = {dr[0], dr[1], dr[2]};
    while (!done) {
        done = 1;
        double r2 = tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2];
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    double s[3];
                    for (int c = 0; c < 3; c++) s[c] = dx * PBC[0][c] + dy * PBC[1][c] + dz * PBC[2][c];
                    tr[0] += s[0]; tr[1] += s[1]; tr[2] += s[2];
                    double nr2 = tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2];
                    if (nr2 >= r2 - 1e-7) {
                        tr[0] -= s[0]; tr[1] -= s[1]; tr[2] -= s[2];
                    } else {
                        r2 = nr2;
                        put[0] += dx; put[1] += dy; put[2] += dz;
                        done = 0;
                    }
                }
    }
    dr[0] = tr[0]; dr[1] = tr[1]; dr[2] = tr[2];
}

static const double KVALS[11] = {0, 0, 0, 0, 0.375, 4.52254248593737e-01, 0.5,
                                 5.30872450464684e-01, 5.51776695296637e-01,
                                 5.66511110779745e-01, 5.77254248593736e-01};
/* the 15 regular Gauss points, weights 1/15 (src/interpLib.C:5336-5375) */
static const double REG_PTS[15][2] = {{0.1, 0.1}, {0.3, 0.1}, {0.5, 0.1}, {0.7, 0.1}, {0.9, 0.1},
                                      {0.1, 0.3}, {0.3, 0.3}, {0.5, 0.3}, {0.7, 0.3}, {0.1, 0.5},
                                      {0.3, 0.5}, {0.5, 0.5}, {0.1, 0.7}, {0.3, 0.7}, {0.1, 0.9}};

#define NREG 15
#define NIRR 16

int orc_plan(omesh *m)
{
    int nv = m->nv;
    int nfr = 0, nfi = 0;
    for (int i = 0; i < nv; i++)
        for (int e = 0; e < m->val[i]; e++) {
            int j = m->nb[i * OMAXV + e], k = m->nb[i * OMAXV + (e + 1) % m->val[i]];
            if (j < i || k < i) continue;
            if (m->val[i] == 6) nfr++;
            else nfi++;
        }
    int nf = nfr + nfi;
    m->nfr = nfr;
    m->nfi = nfi;
    m->face_owner = malloc(sizeof(int) * (size_t)nf);
    m->face_edge = malloc(sizeof(int) * (size_t)nf);
    m->F = calloc((size_t)nfr * NREG + (size_t)nfi * NIRR, sizeof(oform));
    m->face_c0 = calloc((size_t)nf, sizeof(double));

    /* per-valence irregular quadrature: [v_pt][u_pt] -> (u, v, weight) */
    double qu[11][NIRR], qv[11][NIRR], qw[11][NIRR];
    int have_q[11] = {0};

    /* sorted triangle lookup: build a per-vertex list of triangle ids */
    int *vs, *vl;
    vertex_tris(m, m->tri, m->nt, &vs, &vl);

    int fr = 0, fi = 0;
    for (int i = 0; i < nv; i++) {
        int val = m->val[i];
        for (int e = 0; e < val; e++) {
            int ep1 = (e + 1) % val;
            int j = m->nb[i * OMAXV + e], k = m->nb[i * OMAXV + ep1];
            if (j < i || k < i) continue;
            int valj = m->val[j], valk = m->val[k];
            if (val == 6 && (valj != 6 || valk != 6)) { free(vs); free(vl); return -10; }
            int ej = m->rev[i * OMAXV + e], ek = m->rev[i * OMAXV + ep1];
            int pl[OMAXCP];
#define NBI(x) m->nb[i * OMAXV + (((x) % val + val) % val)]
#define NBJ(x) m->nb[j * OMAXV + (((x) % valj + valj) % valj)]
#define NBK(x) m->nb[k * OMAXV + (((x) % valk + valk) % valk)]
            pl[0] = i; pl[1] = k; pl[2] = j;
            pl[3] = NBI(e - 1); pl[4] = NBI(e - 2);
            if (val >= 5) pl[5] = NBI(e - 3);
            if (val >= 6) pl[6] = NBI(e - 4);
            if (val >= 7) pl[7] = NBI(e - 5);
            if (val >= 8) pl[8] = NBI(e - 6);
            pl[val + 1] = NBK(ek - 2);
            pl[val + 2] = NBK(ek + 3);
            pl[val + 3] = NBK(ek + 2);
            pl[val + 4] = NBJ(ej + 3);
            pl[val + 5] = NBJ(ej + 2);
#undef NBI
#undef NBJ
#undef NBK
            int ncp = val + 6;
            /* triangle of this face (sortFaces: faces[e] = tri of {i, nb[e], nb[e+1]}) */
            int s[3] = {i, j, k}, tri = -1;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2 - a; b++)
                    if (s[b] > s[b + 1]) { int t = s[b]; s[b] = s[b + 1]; s[b + 1] = t; }
            for (int x = vs[i]; x < vs[i + 1]; x++) {
                int t = vl[x];
                if (m->tri[3 * t] == s[0] && m->tri[3 * t + 1] == s[1] && m->tri[3 * t + 2] == s[2]) tri = t;
            }
            if (tri < 0) { free(vs); free(vl); return -11; }
            double pbc[3 * OMAXCP];
            for (int p = 0; p < ncp; p++) {
                double dr[3], add[3];
                for (int c = 0; c < 3; c++) dr[c] = m->r[3 * pl[p] + c] - m->r[3 * pl[0] + c];
                min_image(dr, m->PBC, add);
                for (int c = 0; c < 3; c++)
                    pbc[3 * p + c] = add[0] * m->PBC[0][c] + add[1] * m->PBC[1][c] + add[2] * m->PBC[2][c];
            }
            int face, npt;
            oform *F;
            if (val == 6) {
                face = fr++;
                npt = NREG;
                F = m->F + (size_t)face * NREG;
                for (int p = 0; p < NREG; p++) {
                    double B[6][12];
                    orc_basis(REG_PTS[p][0], REG_PTS[p][1], B);
                    for (int d = 0; d < 6; d++)
                        for (int y = 0; y < 12; y++) F[p].t[d][y] = B[d][y];
                    F[p].weight = 1.0 / 15;
                }
            } else {
                face = nfr + fi++;
                npt = NIRR;
                if (val > 10) { free(vs); free(vl); return -12; }
                if (!have_q[val]) {
                    double kk = KVALS[val];
                    double vp[4], wv[4];
                    orc_get_v_pts(vp, wv, 4, kk);
                    int n = 0;
                    for (int a = 0; a < 4; a++) {
                        double fv = vp[a];
                        double v_factor = -(-1 + pow(kk, 4.0 * log2(fv)) * fv * fv * fv * fv * fv) *
                                          log(2.0) / (log(32.0) + 4.0 * log(kk));
                        double v_weight = wv[a] / v_factor;
                        double up[4], wu[4];
                        orc_get_u_pts(up, wu, 4, kk, fv);
                        for (int b = 0; b < 4; b++) {
                            double fu = up[b];
                            double factor = pow(1.0 / (fu + fv) * pow(kk, -log2(fu + fv)), 4.0);
                            qu[val][n] = fu;
                            qv[val][n] = fv;
                            qw[val][n] = v_weight * (wu[b] * factor);
                            n++;
                        }
                    }
                    have_q[val] = 1;
                }
                F = m->F + (size_t)nfr * NREG + (size_t)(face - nfr) * NIRR;
                for (int p = 0; p < NIRR; p++) {
                    double tb[6 * OMAXCP];
                    orc_irregular_table(val, qu[val][p], qv[val][p], tb);
                    for (int d = 0; d < 6; d++)
                        for (int x = 0; x < ncp; x++) F[p].t[d][x] = tb[d * ncp + x];
                    F[p].weight = qw[val][p];
                }
            }
            for (int p = 0; p < npt; p++) {
                F[p].ncp = ncp;
                memcpy(F[p].cp, pl, sizeof(int) * (size_t)ncp);
                memcpy(F[p].pbc, pbc, sizeof(double) * 3 * (size_t)ncp);
                F[p].tri = tri;
            }
            m->face_owner[face] = i;
            m->face_edge[face] = e;
            m->tri_face[tri] = face;
        }
    }
    free(vs);
    free(vl);
    m->planned = 1;
    return 0;
}

void orc_set_params(omesh *m, const oparams *p) { m->prm = *p; }

void orc_set_face_c0(omesh *m, const double *c0)
{
    memcpy(m->face_c0, c0, sizeof(double) * (size_t)(m->nfr + m->nfi));
}

static oform *face_forms(const omesh *m, int f, int *npt, double *dudv)
{
    if (f < m->nfr) {
        *npt = NREG;
        *dudv = 0.5;
        return m->F + (size_t)f * NREG;
    }
    *npt = NIRR;
    *dudv = 1.0;
    return m->F + (size_t)m->nfr * NREG + (size_t)(f - m->nfr) * NIRR;
}

int orc_face_info(const omesh *m, int f, int *owner, int *ncp, int *cp, int *tri)
{
    int npt;
    double dudv;
    oform *F = face_forms(m, f, &npt, &dudv);
    *owner = m->face_owner[f];
    *ncp = F[0].ncp;
    memcpy(cp, F[0].cp, sizeof(int) * (size_t)F[0].ncp);
    *tri = F[0].tri;
    return npt;
}

int orc_face_table(const omesh *m, int f, int p, double *w6, double *weight, double *pbc,
                   double *g0)
{
    int npt;
    double dudv;
    oform *F = face_forms(m, f, &npt, &dudv);
    int ncp = F[p].ncp;
    for (int d = 0; d < 6; d++)
        for (int x = 0; x < ncp; x++) w6[d * ncp + x] = F[p].t[d][x];
    *weight = F[p].weight;
    memcpy(pbc, F[p].pbc, sizeof(double) * 3 * (size_t)ncp);
    *g0 = F[p].g0;
    return ncp;
}

/* total lipid area of each leaflet of triangle t (over the types) */
void orc_tri_lipids(const omesh *m, int t, double *Lo, double *Li)
{
    const int n = m->nlt > 0 ? m->nlt : 1;
    double so = 0.0, si = 0.0;
    for (int x = 0; x < n; x++) {
        so += m->Lo[(size_t)t * n + x];
        si += m->Li[(size_t)t * n + x];
    }
    *Lo = so;
    *Li = si;
}

/* the area of each lipid type in triangle t's leaflets */
void orc_tri_lipid_types(const omesh *m, int t, double *Lo, double *Li)
{
    const int n = m->nlt > 0 ? m->nlt : 1;
    memcpy(Lo, m->Lo + (size_t)t * n, sizeof(double) * (size_t)n);
    memcpy(Li, m->Li + (size_t)t * n, sizeof(double) * (size_t)n);
}

/* glibc's rand() (the TYPE_3 additive feedback generator of random_r:
 * r[i] = r[i-3] + r[i-31] mod 2^32, output r >> 1; srand fills r[0..30]
 * with 16807 r mod (2^31 - 1) from the seed and discards 310 outputs),
 * restated for the reference's rand()-driven lipid packing */
void orc_grand_seed(orc_grand *g, unsigned int seed)
{
    int32_t *r = g->r;
    if (seed == 0) seed = 1;
    int32_t word = (int32_t)seed;   /* as glibc: seeds of 2^31 and above start negative */
    r[0] = word;
    for (int i = 1; i < 31; i++) {
        const int64_t hi = word / 127773, lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        r[i] = word;
    }
    g->f = 3;
    g->b = 0;
    for (int i = 0; i < 310; i++) (void)orc_grand_next(g);
}

int orc_grand_next(orc_grand *g)
{
    const uint32_t val = (uint32_t)g->r[g->f] + (uint32_t)g->r[g->b];
    g->r[g->f] = (int32_t)val;
    g->f = (g->f + 1) % 31;
    g->b = (g->b + 1) % 31;
    return (int)(val >> 1);
}

/* a lipid composition (in place of the single type of the parameters) for
 * the next setg0, whose packing draws from rand() after srand(seed) */
int orc_set_lipids(omesh *m, int n, const double *c0, const double *apl, const double *in_parts,
                   const double *out_parts, unsigned int seed)
{
    if (n < 1) return -1;
    double si = 0.0, so = 0.0;
    for (int x = 0; x < n; x++) {
        if (!(apl[x] > 0.0) || in_parts[x] < 0.0 || out_parts[x] < 0.0) return -1;
        si += in_parts[x];
        so += out_parts[x];
    }
    if (!(si > 0.0) || !(so > 0.0)) return -1;
    free(m->lt_c0); free(m->lt_apl); free(m->lt_in); free(m->lt_out);
    m->lt_c0 = malloc(sizeof(double) * (size_t)n);
    m->lt_apl = malloc(sizeof(double) * (size_t)n);
    m->lt_in = malloc(sizeof(double) * (size_t)n);
    m->lt_out = malloc(sizeof(double) * (size_t)n);
    memcpy(m->lt_c0, c0, sizeof(double) * (size_t)n);
    memcpy(m->lt_apl, apl, sizeof(double) * (size_t)n);
    memcpy(m->lt_in, in_parts, sizeof(double) * (size_t)n);
    memcpy(m->lt_out, out_parts, sizeof(double) * (size_t)n);
    m->nlt = n;
    m->lt_seed = seed;
    free(m->Lo); free(m->Li);
    m->Lo = calloc((size_t)m->nt * (size_t)n, sizeof(double));
    m->Li = calloc((size_t)m->nt * (size_t)n, sizeof(double));
    return 0;
}

/* the lipid types in force: the composition, or the parameters' single type */
static int lipid_types(const omesh *m, const double **c0, const double **apl)
{
    if (m->nlt > 0) {
        *c0 = m->lt_c0;
        *apl = m->lt_apl;
        return m->nlt;
    }
    *c0 = &m->prm.lipid_c0;
    *apl = &m->prm.lipid_apl;
    return 1;
}

/* forward contraction of one Gauss point: X[0..5] = R, Ru, Rv, Ruu, Ruv, Rvv */
static void contract(const omesh *m, const oform *F, const double *r, double X[6][3])
{
    int nv = m->nv;
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    memset(X, 0, sizeof(double) * 18);
    for (int p = 0; p < F->ncp; p++)
        for (int c = 0; c < 3; c++) {
            double x = al[c] * (r[3 * F->cp[p] + c] + F->pbc[3 * p + c]);
            for (int d = 0; d < 6; d++) X[d][c] += F->t[d][p] * x;
        }
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = (a[1] * b[2] - a[2] * b[1]);
    c[1] = -(a[0] * b[2] - a[2] * b[0]);
    c[2] = (a[0] * b[1] - a[1] * b[0]);
}

static void normalize3(double *v)
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (fabs(l) > 0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

/* setg0: reference metric and single-lipid packing (src/interpLib.C:
 * 2969-3388); leaflets: add APL until the packed area exceeds A - 65/2 */
/* Lipid packing of triangle t (several types; src/interpLib.C:3127-3196 for
 * regular faces, 3286-3358 for irregular ones): leaflet fractions from the
 * input parts, cumulated; draws of rand()/RAND_MAX pick a type by the
 * cumulative fractions and add its area until a leaflet holds more than
 * A - 65/2.  Regular faces draw one (inner, outer) pair per round, irregular
 * faces one pair per type within a round. */
static void pack_lipids(omesh *m, orc_grand *gr, int t, double A, int irregular)
{
    const int n = m->nlt;
    double *in = m->Li + (size_t)t * n, *out = m->Lo + (size_t)t * n;
    memset(in, 0, sizeof(double) * (size_t)n);
    memset(out, 0, sizeof(double) * (size_t)n);
    double fr_i[ORC_MAX_LIPID_TYPES], fr_o[ORC_MAX_LIPID_TYPES];
    double sum_i = 0.0, sum_o = 0.0;
    for (int x = 0; x < n; x++) {
        fr_i[x] = m->lt_in[x];
        fr_o[x] = m->lt_out[x];
        sum_i += fr_i[x];
        sum_o += fr_o[x];
    }
    for (int x = 0; x < n; x++) {
        fr_i[x] /= sum_i;
        fr_o[x] /= sum_o;
    }
    for (int x = 1; x < n; x++) fr_i[x] += fr_i[x - 1];
    for (int x = 1; x < n; x++) fr_o[x] += fr_o[x - 1];
    const double area_per_lipid_av = 65.0;
    double a_i = 0.0, a_o = 0.0;
    for (;;) {
        double rn_i = 0.0, rn_o = 0.0;
        if (!irregular) {
            rn_i = orc_grand_next(gr) / 2147483647.0;
            rn_o = orc_grand_next(gr) / 2147483647.0;
        }
        int idone = 0, odone = 0;
        if (a_i > 0 && a_i > A - 0.5 * area_per_lipid_av) idone = 1;
        if (a_o > 0 && a_o > A - 0.5 * area_per_lipid_av) odone = 1;
        if (idone && odone) break;
        for (int x = 0; x < n; x++) {
            if (irregular) {
                rn_i = orc_grand_next(gr) / 2147483647.0;
                rn_o = orc_grand_next(gr) / 2147483647.0;
            }
            if (rn_i < fr_i[x] && !idone) {
                in[x] += m->lt_apl[x];
                idone = 1;
                a_i += m->lt_apl[x];
            }
            if (rn_o < fr_o[x] && !odone) {
                out[x] += m->lt_apl[x];
                odone = 1;
                a_o += m->lt_apl[x];
            }
        }
    }
}

void orc_setg0(omesh *m, const double *r)
{
    int nf = m->nfr + m->nfi;
    double apl = m->prm.lipid_apl;
    orc_grand gr;
    orc_grand_seed(&gr, m->lt_seed);
    for (int f = 0; f < nf; f++) {
        int npt;
        double dudv;
        oform *F = face_forms(m, f, &npt, &dudv);
        double A = 0.0;
        for (int p = 0; p < npt; p++) {
            double X[6][3];
            const double *rr = r;
            /* setg0 contracts without the alpha scale (src/interpLib.C:3047-3075) */
            memset(X, 0, sizeof X);
            for (int q = 0; q < F[p].ncp; q++)
                for (int c = 0; c < 3; c++) {
                    double x = rr[3 * F[p].cp[q] + c] + F[p].pbc[3 * q + c];
                    for (int d = 0; d < 6; d++) X[d][c] += F[p].t[d][q] * x;
                }
            double *Ru = X[1], *Rv = X[2];
            double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            double g = sqrt(RuRu * RvRv - RuRv * RuRv);
            F[p].g0 = g;
            F[p].RuRv0 = RuRv / sqrt(RuRu * RvRv);
            if (f < m->nfr) A += g * dudv * F[p].weight;
            else A += g * F[p].weight;
        }
        int t = F[0].tri;
        if (m->nlt == 0) {
            double a_i = 0.0, a_o = 0.0, Li = 0.0, Lo = 0.0;
            for (;;) {
                int idone = (a_i > 0 && a_i > A - 0.5 * 65.0);
                int odone = (a_o > 0 && a_o > A - 0.5 * 65.0);
                if (idone && odone) break;
                if (!idone) { Li += apl; a_i += apl; }
                if (!odone) { Lo += apl; a_o += apl; }
            }
            m->Li[t] = Li;
            m->Lo[t] = Lo;
        } else {
            pack_lipids(m, &gr, t, A, f >= m->nfr);
        }
    }
    double A0 = 0.0;
    for (int f = 0; f < nf; f++) {
        int npt;
        double dudv;
        oform *F = face_forms(m, f, &npt, &dudv);
        for (int p = 0; p < npt; p++) A0 += dudv * F[p].g0 * F[p].weight;
    }
    m->A0_setg0 = A0;
    /* the reference records V0 right after setg0 on the same positions
     * (src/hd.C:273-279) */
    m->V_setg0 = orc_volume(m, r);
}

/* per-face energy, src/interpLib.C:3426-3881 (mode 0 = FIXED_A +
 * LOCAL_LIPID_ENERGY; mode 1 = per-face c0 Helfrich on the g0 area element) */
double orc_face_energy(omesh *m, int f, const double *r)
{
    int npt;
    double dudv;
    oform *F = face_forms(m, f, &npt, &dudv);
    const oparams *P = &m->prm;
    double e = 0.0;
    for (int p = 0; p < npt; p++) {
        double X[6][3];
        contract(m, F + p, r, X);
        double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
        double nrm[3];
        cross3(Ru, Rv, nrm);
        normalize3(nrm);
        double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
        double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
        double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        double g = sqrt(RuRu * RvRv - RuRv * RuRv);
        double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
        double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
        double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
        double a = 1.0 / (g * g) * (nsuu * RvRv - nsuv * RuRv);
        double b = (1.0 / (g * g)) * (nsuv * RvRv - nsvv * RuRv);
        double c = 1.0 / (g * g) * (nsuv * RuRu - nsuu * RuRv);
        double d = (1.0 / (g * g)) * (nsvv * RuRu - nsuv * RuRv);
        double c1 = -0.5 * (a + d - sqrt(SAFETY_ + a * a + 4 * b * c - 2 * a * d + d * d));
        double c2 = -0.5 * (a + d + sqrt(SAFETY_ + a * a + 4 * b * c - 2 * a * d + d * d));
        double w = F[p].weight, g0 = F[p].g0;
        double en = (P->mode == 0) ? P->kg * c1 * c2
                                   : 0.5 * P->kc * (c1 + c2 - m->face_c0[f]) * (c1 + c2 - m->face_c0[f]) +
                                         P->kg * c1 * c2;
        double dAf = g0 * w;
        double A = dudv * g * w, A0 = dudv * g0 * w;
        double val = RuRv / sqrt(RuRu * RvRv);
        e += P->k_reg * (val - F[p].RuRv0) * (val - F[p].RuRv0);
        if (P->mode == 0) {
            /* every lipid type of both leaflets (src/interpLib.C:3818-3838) */
            int t = F[p].tri;
            const double *lc0, *lapl;
            const int n = lipid_types(m, &lc0, &lapl);
            double lec = 0.0;
            for (int x = 0; x < n; x++) {
                double dc_o = (c1 + c2 - lc0[x]), dc_i = (-c1 - c2 - lc0[x]);
                lec += 0.5 * P->kc * dc_o * dc_o * m->Lo[(size_t)t * n + x] * 0.5 * w;
                lec += 0.5 * P->kc * dc_i * dc_i * m->Li[(size_t)t * n + x] * 0.5 * w;
            }
            e += lec;
        }
        e += dudv * dAf * en;
        e += 2 * 0.5 * A0 * P->KA * ((A - A0) / A0) * ((A - A0) / A0);
    }
    return e;
}

/* The face's curvature-energy density as fenergy / ifenergy form it for
 * penergy (src/interpLib.C:3850-3851 regular, 3612-3613 irregular, FIXED_A
 * so dAf = g0 w): regular faces sum dudv dAf 1/2 kc (c1 + c2 - c0)^2, irregular
 * faces dudv dAf en with en the point energy density (kg c1 c2 under
 * LOCAL_LIPID_ENERGY, else 1/2 kc (c1 + c2 - c0)^2 + kg c1 c2); area the sum
 * of dudv dAf. */
int orc_face_density(omesh *m, int f, const double *r, double *num, double *area)
{
    int npt;
    double dudv;
    if (f < 0 || f >= m->nfr + m->nfi) return -1;
    oform *F = face_forms(m, f, &npt, &dudv);
    const oparams *P = &m->prm;
    const double c0 = m->face_c0[f];
    double s = 0.0, a = 0.0;
    for (int p = 0; p < npt; p++) {
        double X[6][3];
        contract(m, F + p, r, X);
        double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
        double nrm[3];
        cross3(Ru, Rv, nrm);
        normalize3(nrm);
        double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
        double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
        double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        double g = sqrt(RuRu * RvRv - RuRv * RuRv);
        double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
        double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
        double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
        double sa = 1.0 / (g * g) * (nsuu * RvRv - nsuv * RuRv);
        double sb = (1.0 / (g * g)) * (nsuv * RvRv - nsvv * RuRv);
        double sc = 1.0 / (g * g) * (nsuv * RuRu - nsuu * RuRv);
        double sd = (1.0 / (g * g)) * (nsvv * RuRu - nsuv * RuRv);
        double c1 = -0.5 * (sa + sd - sqrt(SAFETY_ + sa * sa + 4 * sb * sc - 2 * sa * sd + sd * sd));
        double c2 = -0.5 * (sa + sd + sqrt(SAFETY_ + sa * sa + 4 * sb * sc - 2 * sa * sd + sd * sd));
        double dAf = F[p].g0 * F[p].weight;
        double x;
        if (f < m->nfr)
            x = 0.5 * P->kc * (c1 + c2 - c0) * (c1 + c2 - c0);
        else
            x = (P->mode == 0) ? P->kg * c1 * c2 : 0.5 * P->kc * (c1 + c2 - c0) * (c1 + c2 - c0) + P->kg * c1 * c2;
        s += dudv * dAf * x;
        a += dudv * dAf;
    }
    *num = s;
    *area = a;
    return 0;
}

/* surface::penergy for one particle (src/interpLib.C:8330-8790, do_monge 0):
 * 1/2 kc pa (c1 + c2 - pc0)^2 - pa fed, c1 and c2 at the particle's (u, v)
 * (the curvatures do not change with the sub-patch's power-of-two scale, so
 * the point geometry's serve), fed the face's density over its area */
int orc_particle_energy(const omesh *m, const double *r, int f, double u, double v, double pc0, double pa,
                        double fed, double *out)
{
    double gq[15];
    if (orc_point_geometry(m, r, f, u, v, gq) != 0) return -1;
    const double c1 = gq[9], c2 = gq[10];
    double en = 0.0;
    en += 0.5 * m->prm.kc * pa * (c1 + c2 - pc0) * (c1 + c2 - pc0);
    en -= pa * fed;
    *out = en;
    return 0;
}

/* surface::energy: faces summed in triangle order (src/interpLib.C:4315-4320,
 * src/parallel.C:319-326) */
double orc_energy(omesh *m, const double *r)
{
    double e = 0.0;
    for (int t = 0; t < m->nt; t++) {
        int f = m->tri_face[t];
        if (f >= 0) e += orc_face_energy(m, f, r);
    }
    return e + orc_restraint_energy(m, r, NULL, NULL);
}

/* sum over the regular faces' Gauss points of 1/2 kc (c1+c2-c0)^2 dudv w g(r)
 * with c1 + c2 from r_star: its gradient at r = r_star is the dE/dg term
 * that surface::grad adds in the Helfrich form (src/grad.C:448).  Used only
 * to pin the gradient by finite differences in r with the curvature frozen. */
double orc_frozen_curvature_area(omesh *m, const double *r_star, const double *r)
{
    const oparams *P = &m->prm;
    double s = 0.0;
    for (int f = 0; f < m->nfr; f++) {
        int npt;
        double dudv;
        oform *F = face_forms(m, f, &npt, &dudv);
        const double c0 = m->face_c0[f];
        for (int p = 0; p < npt; p++) {
            double X[6][3], Y[6][3];
            contract(m, F + p, r_star, X);
            contract(m, F + p, r, Y);
            double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
            double nrm[3];
            cross3(Ru, Rv, nrm);
            normalize3(nrm);
            double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            double det = RuRu * RvRv - RuRv * RuRv;
            double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
            double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
            double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
            /* c1 + c2 = -tr(I^-1 II) */
            double H = -(nsuu * RvRv - 2.0 * nsuv * RuRv + nsvv * RuRu) / det;
            double *Su = Y[1], *Sv = Y[2];
            double n[3];
            cross3(Su, Sv, n);
            double g = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            s += 0.5 * P->kc * (H - c0) * (H - c0) * dudv * F[p].weight * g;
        }
    }
    return s;
}

/* one Gauss point of grad/igrad (src/grad.C:246-930, 6605-7450), through the
 * reference's graph: E(c1, c2, g) -> (a,b,c,d,g) -> (nsuu,nsuv,nsvv,RuRu,RuRv,
 * RvRv) -> (Ru, Rv, Suu, Suv, Svv) -> control points */
static void point_grad(omesh *m, int f, const oform *F, double dudv, const double *r, double *gr,
                       double lamA)
{
    const oparams *P = &m->prm;
    int nv = m->nv;
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    double X[6][3];
    contract(m, F, r, X);
    double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
    double nrm[3];
    cross3(Ru, Rv, nrm);
    normalize3(nrm);
    double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    double g = sqrt(RuRu * RvRv - RuRv * RuRv);
    double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
    double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
    double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
    double g2 = g * g, g3 = g * g * g;
    double a = 1.0 / g2 * (nsuu * RvRv - nsuv * RuRv);
    double b = (1.0 / g2) * (nsuv * RvRv - nsvv * RuRv);
    double c = 1.0 / g2 * (nsuv * RuRu - nsuu * RuRv);
    double d = (1.0 / g2) * (nsvv * RuRu - nsuv * RuRv);
    double S = sqrt(SAFETY_ + a * a + 4 * b * c - 2 * a * d + d * d);
    double e1 = -0.5 * (a + d - S), e2 = -0.5 * (a + d + S);
    double w = F->weight, g0 = F->g0;
    double dAf = g0 * w;
    double A = dudv * g * w, A0 = dudv * g0 * w;

    /* dE/dc1, dE/dc2, dE/dg */
    double de_c1 = 0.0, de_c2 = 0.0, de_g = 0.0;
    if (P->mode == 0) {
        /* every lipid type (src/grad.C:405-423) */
        int t = F->tri;
        const double *lc0, *lapl;
        const int n = lipid_types(m, &lc0, &lapl);
        for (int x = 0; x < n; x++) {
            double dc_o = (e1 + e2 - lc0[x]), dc_i = (-e1 - e2 - lc0[x]);
            de_c1 += P->kc * dc_o * 0.5 * m->Lo[(size_t)t * n + x] * w;
            de_c1 -= P->kc * dc_i * 0.5 * m->Li[(size_t)t * n + x] * w;
            de_c2 += P->kc * dc_o * 0.5 * m->Lo[(size_t)t * n + x] * w;
            de_c2 -= P->kc * dc_i * 0.5 * m->Li[(size_t)t * n + x] * w;
        }
    } else {
        double c0 = m->face_c0[f];
        de_c1 += (P->kc * (e1 + e2 - c0)) * dudv * dAf;
        de_c2 += (P->kc * (e1 + e2 - c0)) * dudv * dAf;
    }
    de_c1 += P->kg * e2 * dudv * dAf;
    de_c2 += P->kg * e1 * dudv * dAf;
    /* Helfrich form, regular faces: surface::grad adds the area element's
     * derivative of 1/2 kc (c1+c2-c0)^2 dudv g w although the energy holds g0
     * fixed (the #else of LOCAL_LIPID_ENERGY, src/grad.C:448, outside its
     * #ifndef FIXED_A); igrad keeps the same line inside #ifndef FIXED_A
     * (src/grad.C:6822-6824), so irregular faces do not add it */
    if (P->mode != 0 && f < m->nfr && P->grad_form == 0) {
        double c0 = m->face_c0[f];
        de_g += (0.5 * P->kc * (e1 + e2 - c0) * (e1 + e2 - c0)) * dudv * w;
    }
    de_g += 2 * P->KA * ((A - A0) / A0) * w * dudv;
    /* global area restraint: dE/dA_tot * dA_tot/dg, A_tot = sum dudv g w */
    de_g += lamA * w * dudv;

    /* c1, c2 as functions of the shape operator entries (src/grad.C:381-399) */
    double dc1_a = -0.5 * (1 - (2 * a - 2 * d) / (2. * S));
    double dc1_b = c / S, dc1_c = b / S;
    double dc1_d = -0.5 * (1 - (-2 * a + 2 * d) / (2. * S));
    double dc2_a = -0.5 * (1 + (2 * a - 2 * d) / (2. * S));
    double dc2_b = -c / S, dc2_c = -b / S;
    double dc2_d = -0.5 * (1 + (-2 * a + 2 * d) / (2. * S));
    double de_a = de_c1 * dc1_a + de_c2 * dc2_a;
    double de_b = de_c1 * dc1_b + de_c2 * dc2_b;
    double de_c = de_c1 * dc1_c + de_c2 * dc2_c;
    double de_d = de_c1 * dc1_d + de_c2 * dc2_d;

    /* shape operator entries as functions of (g, II, I) (src/grad.C:364-386) */
    de_g += de_a * (-2 * (-(nsuv * RuRv) + nsuu * RvRv) / g3);
    de_g += de_b * (-2 * (-(nsvv * RuRv) + nsuv * RvRv) / g3);
    de_g += de_c * (-2 * (nsuv * RuRu - nsuu * RuRv) / g3);
    de_g += de_d * (-2 * (nsvv * RuRu - nsuv * RuRv) / g3);
    double de_nsuu = de_a * (RvRv / g2) + de_c * (-RuRv / g2);
    double de_nsuv = de_a * (-RuRv / g2) + de_b * (RvRv / g2) + de_c * (RuRu / g2) + de_d * (-RuRv / g2);
    double de_nsvv = de_b * (-RuRv / g2) + de_d * (RuRu / g2);
    double de_RuRu = de_c * (nsuv / g2) + de_d * (nsvv / g2);
    double de_RvRv = de_a * (nsuu / g2) + de_b * (nsuv / g2);
    double de_RuRv = de_a * (-nsuv / g2) + de_b * (-nsvv / g2) + de_c * (-nsuu / g2) + de_d * (-nsuv / g2);
    /* g(RuRu, RuRv, RvRv) (src/grad.C:344-346) */
    de_RuRu += de_g * RvRv / (2. * sqrt(-RuRv * RuRv + RuRu * RvRv));
    de_RuRv += de_g * (-(RuRv / sqrt(-RuRv * RuRv + RuRu * RvRv)));
    de_RvRv += de_g * RuRu / (2. * sqrt(-RuRv * RuRv + RuRu * RvRv));

    double dE[5][3]; /* dE/dRu, dE/dRv, dE/dSuu, dE/dSuv, dE/dSvv */
    double de_n[3];
    for (int k = 0; k < 3; k++) {
        de_n[k] = de_nsuu * Suu[k] + de_nsuv * Suv[k] + de_nsvv * Svv[k];
        dE[2][k] = 0.0;
        dE[0][k] = 2 * Ru[k] * de_RuRu + Rv[k] * de_RuRv;
        dE[1][k] = 2 * Rv[k] * de_RvRv + Ru[k] * de_RuRv;
        dE[2][k] = de_nsuu * nrm[k];
        dE[3][k] = de_nsuv * nrm[k];
        dE[4][k] = de_nsvv * nrm[k];
    }
    /* d nrm / d(Ru, Rv): nrm = N/|N|, N = Ru x Rv (src/grad.C:470-494) */
    {
        double N[3];
        cross3(Ru, Rv, N);
        double n2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
        double fac = pow(n2, 1.5);
        double J[3][3]; /* dn_i/dN_j */
        for (int i2 = 0; i2 < 3; i2++)
            for (int j2 = 0; j2 < 3; j2++) J[i2][j2] = ((i2 == j2) * n2 - N[i2] * N[j2]) / fac;
        double de_N[3];
        for (int j2 = 0; j2 < 3; j2++)
            de_N[j2] = de_n[0] * J[0][j2] + de_n[1] * J[1][j2] + de_n[2] * J[2][j2];
        /* dN = dRu x Rv + Ru x dRv */
        double t1[3], t2[3];
        cross3(Rv, de_N, t1); /* d/dRu of de_N . (Ru x Rv) = Rv x de_N */
        cross3(de_N, Ru, t2); /* d/dRv = de_N x Ru */
        for (int k = 0; k < 3; k++) {
            dE[0][k] += t1[k];
            dE[1][k] += t2[k];
        }
    }
    /* k_reg term on val = RuRv/sqrt(RuRu RvRv) (src/grad.C:855-861) */
    double val = RuRv / sqrt(RuRu * RvRv);
    double dval_du[3], dval_dv[3];
    for (int k = 0; k < 3; k++) {
        dval_du[k] = Rv[k] / sqrt(RuRu * RvRv) - 2 * Ru[k] * RvRv * RuRv / (2 * pow(RuRu * RvRv, 1.5));
        dval_dv[k] = Ru[k] / sqrt(RuRu * RvRv) - 2 * Rv[k] * RuRu * RuRv / (2 * pow(RuRu * RvRv, 1.5));
    }
    double kr = P->k_reg * 2 * (val - F->RuRv0);
    /* scatter (src/grad.C:863-930) */
    for (int p = 0; p < F->ncp; p++) {
        int v = F->cp[p];
        double dedr[3];
        for (int k = 0; k < 3; k++) {
            double s = dE[0][k] * F->t[1][p];
            s += dE[1][k] * F->t[2][p];
            s += dE[2][k] * F->t[3][p];
            s += dE[3][k] * F->t[4][p];
            s += dE[4][k] * F->t[5][p];
            s += kr * dval_du[k] * F->t[1][p];
            s += kr * dval_dv[k] * F->t[2][p];
            dedr[k] = s;
            gr[3 * v + k] += s * al[k];
        }
        for (int k = 0; k < 3; k++) gr[3 * nv + k] += dedr[k] * (r[3 * v + k] + F->pbc[3 * p + k]);
    }
}

void orc_grad(omesh *m, const double *r, double *g)
{
    double lamA = 0.0, lamV = 0.0;
    orc_restraint_energy(m, r, &lamA, &lamV);
    /* regular faces in triangle order, then irregular (grad then igrad) */
    for (int pass = 0; pass < 2; pass++)
        for (int t = 0; t < m->nt; t++) {
            int f = m->tri_face[t];
            if (f < 0) continue;
            if ((pass == 0) != (f < m->nfr)) continue;
            int npt;
            double dudv;
            oform *F = face_forms(m, f, &npt, &dudv);
            for (int p = 0; p < npt; p++) point_grad(m, f, F + p, dudv, r, g, lamA);
        }
    if (lamV != 0.0) {
        /* volume restraint: KV strain dV/dr (src/minimize.C:351-376) */
        int n3 = 3 * m->nv;
        double *dV = malloc(sizeof(double) * (size_t)n3);
        orc_volume_grad(m, r, dV);
        for (int i = 0; i < n3; i++) g[i] += lamV * dV[i];
        free(dV);
    }
}

/* Energy and gradient with the faces split over nthreads OpenMP threads,
 * each with its own gradient buffer, summed in thread order afterwards (the
 * multi-core CPU comparator; the serial functions above are the reference's
 * single-threaded path).  Same terms as orc_energy + orc_grad; g is
 * overwritten. */
double orc_energy_grad_threads(omesh *m, const double *r, double *g, int nthreads)
{
    double lamA = 0.0, lamV = 0.0;
    const double er = orc_restraint_energy(m, r, &lamA, &lamV);
    const int n3 = 3 * m->nv;
    if (nthreads < 1) nthreads = 1;
    double *gbuf = calloc((size_t)nthreads * (size_t)(n3 + 3), sizeof(double));
    double e = 0.0;
#pragma omp parallel num_threads(nthreads) reduction(+ : e)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double *gt = gbuf + (size_t)tid * (size_t)(n3 + 3);
#pragma omp for schedule(static)
        for (int t = 0; t < m->nt; t++) {
            int f = m->tri_face[t];
            if (f < 0) continue;
            e += orc_face_energy(m, f, r);
            int npt;
            double dudv;
            oform *F = face_forms(m, f, &npt, &dudv);
            for (int p = 0; p < npt; p++) point_grad(m, f, F + p, dudv, r, gt, lamA);
        }
    }
    memset(g, 0, sizeof(double) * (size_t)(n3 + 3));
    for (int k = 0; k < nthreads; k++)
        for (int i = 0; i < n3 + 3; i++) g[i] += gbuf[(size_t)k * (size_t)(n3 + 3) + i];
    free(gbuf);
    if (lamV != 0.0) {
        double *dV = malloc(sizeof(double) * (size_t)n3);
        orc_volume_grad(m, r, dV);
        for (int i = 0; i < n3; i++) g[i] += lamV * dV[i];
        free(dV);
    }
    return e + er;
}

/* geometric area sum(dudv g w) and reference area sum(dudv g0 w) */
void orc_area(omesh *m, const double *r, double *A, double *A0)
{
    *A = 0.0;
    *A0 = 0.0;
    for (int t = 0; t < m->nt; t++) {
        int f = m->tri_face[t];
        if (f < 0) continue;
        int npt;
        double dudv;
        oform *F = face_forms(m, f, &npt, &dudv);
        for (int p = 0; p < npt; p++) {
            double X[6][3];
            contract(m, F + p, r, X);
            double *Ru = X[1], *Rv = X[2];
            double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            double g = sqrt(RuRu * RvRv - RuRv * RuRv);
            *A += dudv * g * F[p].weight;
            *A0 += dudv * F[p].g0 * F[p].weight;
        }
    }
}

/* ---- global restraints ------------------------------------------------ */

void orc_set_restraints(omesh *m, double KA_global, double KV, double A0, double V0)
{
    m->KA_global = KA_global;
    m->KV = KV;
    m->A0_target = A0;
    m->V0_target = V0;
}

/* Limit position of every vertex.  The reference's volume plan evaluates the
 * limit surface at the corners of each face (ni = 2, src/interpLib.C:307,
 * 426-461): the top eigenvector of the valence-n subdivision stencil at the
 * owner corner, the box spline (or the Stam route) at the other two.  With the
 * stencil weight beta = 3/(8n) (src/interpLib.C:134) every one of these is
 * Loop's limit mask: L_v = (1 - n chi) r_v + chi sum(ring), chi = 1/(n +
 * 3/(8 beta)) = 1/(2n).  Periodic meshes take the minimum image of each ring
 * vertex relative to v. */
void orc_limit_points(const omesh *m, const double *r, double *L)
{
    for (int v = 0; v < m->nv; v++) {
        int n = m->val[v];
        double s[3] = {0.0, 0.0, 0.0};
        for (int e = 0; e < n; e++) {
            int u = m->nb[v * OMAXV + e];
            double dr[3], add[3];
            for (int c = 0; c < 3; c++) dr[c] = r[3 * u + c] - r[3 * v + c];
            min_image(dr, (double (*)[3])m->PBC, add);
            for (int c = 0; c < 3; c++)
                s[c] += r[3 * u + c] + add[0] * m->PBC[0][c] + add[1] * m->PBC[1][c] + add[2] * m->PBC[2][c];
        }
        double chi = 1.0 / (2.0 * n);
        for (int c = 0; c < 3; c++) L[3 * v + c] = (1.0 - n * chi) * r[3 * v + c] + chi * s[c];
    }
}

/* Enclosed volume of the polyhedron of limit points, one flat triangle per
 * face (surface::volume at ni = 2, src/interpLib.C:997-1098; the in-box branch
 * of new_volume, src/volume.C:350-430): r1 = owner corner (u,v) = (0,0), r2 =
 * the u = 1 corner (ring[e]), r3 = the v = 1 corner (ring[e+1]); vol +=
 * (r1 . n) A / 3 with n, A the unit normal and area of (r2 - r1) x (r3 - r1).
 * Faces in plan order (owner ascending, ring position ascending).  The
 * reference's box-centre shift cancels for a closed surface and is omitted. */
double orc_volume(const omesh *m, const double *r)
{
    double *L = malloc(sizeof(double) * 3 * (size_t)m->nv);
    orc_limit_points(m, r, L);
    double vol = 0.0;
    for (int i = 0; i < m->nv; i++) {
        int val = m->val[i];
        for (int e = 0; e < val; e++) {
            int j = m->nb[i * OMAXV + e], k = m->nb[i * OMAXV + (e + 1) % val];
            if (j < i || k < i) continue;
            const double *r1 = L + 3 * i, *r2 = L + 3 * j, *r3 = L + 3 * k;
            double d1[3] = {r2[0] - r1[0], r2[1] - r1[1], r2[2] - r1[2]};
            double d2[3] = {r3[0] - r1[0], r3[1] - r1[1], r3[2] - r1[2]};
            double nrm[3];
            cross3(d1, d2, nrm);
            double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            double A = len / 2;
            for (int c = 0; c < 3; c++) nrm[c] /= len;
            double kk = r1[0] * nrm[0] + r1[1] * nrm[1] + r1[2] * nrm[2];
            vol += kk * A / 3.0;
        }
    }
    free(L);
    return vol;
}

/* dV/dr (3 nv, overwritten): per face dV/dL_i = (L_j x L_k)/6 and cyclic
 * (the closed form of the derivative src/volume.C:258-430 builds term by
 * term), then through the limit mask: dV/dr_u = sum_v M_vu dV/dL_v */
void orc_volume_grad(const omesh *m, const double *r, double *dV)
{
    int nv = m->nv;
    double *L = malloc(sizeof(double) * 3 * (size_t)nv);
    double *D = calloc(3 * (size_t)nv, sizeof(double));
    orc_limit_points(m, r, L);
    for (int i = 0; i < nv; i++) {
        int val = m->val[i];
        for (int e = 0; e < val; e++) {
            int j = m->nb[i * OMAXV + e], k = m->nb[i * OMAXV + (e + 1) % val];
            if (j < i || k < i) continue;
            double t[3];
            cross3(L + 3 * j, L + 3 * k, t);
            for (int c = 0; c < 3; c++) D[3 * i + c] += t[c] / 6.0;
            cross3(L + 3 * k, L + 3 * i, t);
            for (int c = 0; c < 3; c++) D[3 * j + c] += t[c] / 6.0;
            cross3(L + 3 * i, L + 3 * j, t);
            for (int c = 0; c < 3; c++) D[3 * k + c] += t[c] / 6.0;
        }
    }
    memset(dV, 0, sizeof(double) * 3 * (size_t)nv);
    for (int v = 0; v < nv; v++) {
        int n = m->val[v];
        double chi = 1.0 / (2.0 * n);
        for (int c = 0; c < 3; c++) dV[3 * v + c] += (1.0 - n * chi) * D[3 * v + c];
        for (int e = 0; e < n; e++) {
            int u = m->nb[v * OMAXV + e];
            for (int c = 0; c < 3; c++) dV[3 * u + c] += chi * D[3 * v + c];
        }
    }
    free(L);
    free(D);
}

/* Restraint energy at r and the multipliers dE/dA_tot, dE/dV:
 *   area   KA_g A0 ((A - A0)/A0)^2, A = sum dudv g w (src/interpLib.C:
 *          4322-4326, the GLOBAL_AREA branch; A0 = sum dudv g0 w unless set)
 *   volume KV/2 V0 ((V - V0)/V0)^2 (src/minimize.C:90-114; V0 = V at setg0
 *          unless set, src/hd.C:273-279) */
double orc_restraint_energy(omesh *m, const double *r, double *lamA, double *lamV)
{
    double e = 0.0;
    if (lamA) *lamA = 0.0;
    if (lamV) *lamV = 0.0;
    if (m->KA_global != 0.0) {
        double A, A0g;
        orc_area(m, r, &A, &A0g);
        double A0 = m->A0_target > 0.0 ? m->A0_target : m->A0_setg0;
        double s = (A - A0) / A0;
        e += m->KA_global * A0 * s * s;
        if (lamA) *lamA = 2.0 * m->KA_global * s;
    }
    if (m->KV != 0.0) {
        double V = orc_volume(m, r);
        double V0 = m->V0_target != 0.0 ? m->V0_target : m->V_setg0;
        double s = (V - V0) / V0;
        e += 0.5 * m->KV * V0 * s * s;
        if (lamV) *lamV = m->KV * s;
    }
    return e;
}

/* nsteps of r <- r - mobility_dt * dE/dr + noise_amp * xi_s; energies[s] is
 * the energy at the start of step s (may be NULL) */
void orc_bd_steps(omesh *m, double *r, const double *xi, int nsteps, double mobility_dt,
                  double noise_amp, double *energies)
{
    int n3 = 3 * m->nv;
    double *g = malloc(sizeof(double) * (size_t)(n3 + 3));
    for (int s = 0; s < nsteps; s++) {
        if (energies) energies[s] = orc_energy(m, r);
        memset(g, 0, sizeof(double) * (size_t)(n3 + 3));
        orc_grad(m, r, g);
        const double *x = xi + (size_t)s * n3;
        for (int i = 0; i < n3; i++) r[i] += -mobility_dt * g[i] + noise_amp * x[i];
    }
    free(g);
}

/* the same steps with each energy + gradient split over nthreads OpenMP
 * threads (orc_energy_grad_threads): the threaded CPU comparator */
void orc_bd_steps_threads(omesh *m, double *r, const double *xi, int nsteps, double mobility_dt,
                          double noise_amp, double *energies, int nthreads)
{
    int n3 = 3 * m->nv;
    double *g = malloc(sizeof(double) * (size_t)(n3 + 3));
    for (int s = 0; s < nsteps; s++) {
        const double e = orc_energy_grad_threads(m, r, g, nthreads);
        if (energies) energies[s] = e;
        const double *x = xi + (size_t)s * n3;
        for (int i = 0; i < n3; i++) r[i] += -mobility_dt * g[i] + noise_amp * x[i];
    }
    free(g);
}

/* ---- effective-mass integrators (src/hd.C:418-481, 1543-1720, 1943-1947; */
/* src/modes.C:1719-1960; src/uv_map.C:14-200; src/parallel.C:1028-1075)    */

#define ORC_AMU_PER_KG (1.0 / 1.661e-27)          /* src/units.h */
#define ORC_AKMA_TIME (1.0 / (0.04888 * (1e-12)))  /* src/units.h */
#define ORC_PLIM 10                                /* src/hd.C:419 */

int orc_effm_points_per_face(void) { return (ORC_PLIM + 1) * (ORC_PLIM + 2) / 2; }

/* mass of one of the ppf points on each face (src/hd.C:449-455): two leaflets
 * of POPC (MSCALE 5) over the reference area area0 */
double orc_effm_point_mass(omesh *m)
{
    double A, A0;
    double *r = malloc(sizeof(double) * (3 * (size_t)m->nv + 3));
    orc_get_positions(m, r);
    orc_area(m, r, &A, &A0);
    free(r);
    const double MSCALE = 5;
    const double mass_per_lipid = MSCALE * 1000 * 760.09 / 6.022e23 / 1000;
    const double area_per_lipid = 65.35;
    const int ppf = orc_effm_points_per_face();
    return ORC_AMU_PER_KG * 2 * A0 / (m->nfr + m->nfi) / ppf * mass_per_lipid / area_per_lipid;
}

/* position coefficients of the limit surface at (u, v) of face f
 * (surface::get_pt_coeffs, src/uv_map.C:14-200): the box spline on regular
 * faces, the subdivision route on irregular ones, and at the extraordinary
 * vertex itself (u = v = 0) its limit mask (1/2 on the vertex, 1/(2 val) on
 * each neighbour, src/uv_map.C:36-64).  Returns the number of coefficients. */
static int pt_coeffs(const omesh *m, int f, double u, double v, double *coef, int *list)
{
    int owner, ncp, cp[OMAXCP], tri;
    orc_face_info(m, f, &owner, &ncp, cp, &tri);
    if (f >= m->nfr) {
        int val = m->val[owner];
        if (u < 1e-10 && v < 1e-10) {
            coef[0] = 0.5;
            list[0] = owner;
            for (int e = 0; e < val; e++) {
                coef[1 + e] = 1.0 / (val * 2);
                list[1 + e] = m->nb[owner * OMAXV + e];
            }
            return 1 + val;
        }
        double *tab = malloc(sizeof(double) * 6 * (size_t)ncp);
        orc_irregular_table(val, u, v, tab);
        for (int x = 0; x < ncp; x++) {
            coef[x] = tab[x];
            list[x] = cp[x];
        }
        free(tab);
        return ncp;
    }
    double B[6][12];
    orc_basis(u, v, B);
    for (int x = 0; x < ncp; x++) {
        coef[x] = B[0][x];
        list[x] = cp[x];
    }
    return ncp;
}

/* M_ij = sum over triangles (src/hd.C:459-478 triangle order) and the
 * (plim+1)(plim+2)/2 lattice points (i/plim, j/plim) of m c_i c_j
 * (src/modes.C:1736-1741); M is nv x nv, row-major */
void orc_effm_mass(omesh *m, double mass_per_pt, double *M)
{
    int nv = m->nv;
    memset(M, 0, sizeof(double) * (size_t)nv * nv);
    double coef[OMAXCP];
    int list[OMAXCP];
    for (int t = 0; t < m->nt; t++) {
        int f = m->tri_face[t];
        if (f < 0) continue;
        for (int fi = 0; fi <= ORC_PLIM; fi++)
            for (int fj = 0; fj <= ORC_PLIM - fi; fj++) {
                int n = pt_coeffs(m, f, fi / (double)ORC_PLIM, fj / (double)ORC_PLIM, coef, list);
                for (int c1 = 0; c1 < n; c1++)
                    for (int c2 = 0; c2 < n; c2++)
                        M[(size_t)list[c1] * nv + list[c2]] += coef[c1] * coef[c2] * mass_per_pt;
            }
    }
}

/* in-place inverse of a symmetric positive-definite matrix by Cholesky
 * (the dpotrf + dpotri of src/modes.C:1827-1847, textbook form: A = L L^T,
 * A^-1 = L^-T L^-1); returns 0, or the failing pivot + 1 */
int orc_effm_invert(int n, double *A)
{
    double *L = calloc((size_t)n * n, sizeof(double));
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(d > 0.0)) {
            free(L);
            return j + 1;
        }
        double ljj = sqrt(d);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = s / ljj;
        }
    }
    /* W = L^-1 (lower triangular), column by column */
    double *W = calloc((size_t)n * n, sizeof(double));
    for (int j = 0; j < n; j++) {
        W[(size_t)j * n + j] = 1.0 / L[(size_t)j * n + j];
        for (int i = j + 1; i < n; i++) {
            double s = 0.0;
            for (int k = j; k < i; k++) s -= L[(size_t)i * n + k] * W[(size_t)k * n + j];
            W[(size_t)i * n + j] = s / L[(size_t)i * n + i];
        }
    }
    /* A^-1 = W^T W */
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
            double s = 0.0;
            for (int k = i; k < n; k++) s += W[(size_t)k * n + i] * W[(size_t)k * n + j];
            A[(size_t)i * n + j] = A[(size_t)j * n + i] = s;
        }
    free(L);
    free(W);
    return 0;
}

/* truncation of the inverse (src/modes.C:1889-1905, rel_tol = 1e-20 at
 * src/modes.C:16): entries with |Mi_ij| / sqrt(|Mi_ii Mi_jj|) <= rel_tol are
 * dropped (set to zero); returns the number kept */
long orc_effm_truncate(int n, double *Mi, double rel_tol)
{
    long kept = 0;
    double *d = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) d[i] = Mi[(size_t)i * n + i];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double *x = Mi + (size_t)i * n + j;
            if (fabs(*x) / sqrt(fabs(d[i] * d[j])) > rel_tol) kept++;
            else *x = 0.0;
        }
    free(d);
    return kept;
}

/* vout[v][c] = (Mi . vin_c)[v] (src/parallel.C:1028-1075 without the scale) */
static void effm_mult(int nv, const double *Mi, const double *vin, double *vout)
{
    for (int v = 0; v < nv; v++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int u = 0; u < nv; u++) s += Mi[(size_t)v * nv + u] * vin[3 * u + c];
            vout[3 * v + c] = s;
        }
}

/* nsteps of the Cartesian momentum-space integrator of src/hd.C with the
 * inverse effective mass Mi (nv x nv): mode 0 = do_bd_membrane (momenta reset
 * after every step, src/hd.C:1943-1947), mode 1 = do_ld with OLD_LANGEVIN
 * friction through Mi (src/hd.C:40, 1546-1557).  Per step: g = grad(r);
 * friction (mode 1); half kick; noise (mode 0: gamma^-1 N(0, s) / Delta,
 * mode 1: N(0, s), s = sqrt(2 gamma kT Delta), xi[step][3nv] in Q = 3v + c
 * order, src/hd.C:1654-1673); half kick; r += (Mi pp) / alpha^2 Delta
 * (src/hd.C:1699-1720).  pp[3nv] carries the momenta (mode 1). */
void orc_effm_steps(omesh *m, double *r, double *pp, const double *xi, int nsteps, int mode, double dt,
                    double kT, double gamma, const double *Mi, double *energies)
{
    int nv = m->nv, n3 = 3 * nv;
    double *g = malloc(sizeof(double) * (size_t)(n3 + 3));
    double *next = malloc(sizeof(double) * (size_t)n3);
    double *fr = malloc(sizeof(double) * (size_t)n3);
    const double akma = ORC_AKMA_TIME;
    for (int s = 0; s < nsteps; s++) {
        if (energies) energies[s] = orc_energy(m, r);
        memset(g, 0, sizeof(double) * (size_t)(n3 + 3));
        orc_grad(m, r, g);
        const double al[3] = {r[n3], r[n3 + 1], r[n3 + 2]};
        if (mode == 0) memset(pp, 0, sizeof(double) * (size_t)n3);
        memcpy(next, pp, sizeof(double) * (size_t)n3);
        if (mode == 1) {
            const double scale = -akma * dt;
            effm_mult(nv, Mi, pp, fr);
            for (int v = 0; v < nv; v++)
                for (int c = 0; c < 3; c++) next[3 * v + c] += fr[3 * v + c] * (gamma * scale / (al[c] * al[c]));
        }
        for (int i = 0; i < n3; i++) next[i] += -g[i] * akma * dt / 2;
        memcpy(pp, next, sizeof(double) * (size_t)n3);
        const double *x = xi + (size_t)s * n3;
        const double sig = sqrt(2 * gamma * kT * akma * dt);
        for (int i = 0; i < n3; i++) {
            if (mode == 0) next[i] += (1.0 / gamma) * (sig * x[i]) / (akma * dt);
            else next[i] += sig * x[i];
        }
        for (int i = 0; i < n3; i++) next[i] += -g[i] * akma * dt / 2;
        memcpy(pp, next, sizeof(double) * (size_t)n3);
        effm_mult(nv, Mi, pp, fr);
        for (int v = 0; v < nv; v++)
            for (int c = 0; c < 3; c++) r[3 * v + c] += fr[3 * v + c] * (1.0 / (al[c] * al[c])) * akma * dt;
        if (mode == 0) memset(pp, 0, sizeof(double) * (size_t)n3);
    }
    free(g);
    free(next);
    free(fr);
}

/* ------------------------------------------------------------------------ */
/* Points on the limit surface: surface::evaluateRNRM (src/interpLib.C:     */
/* 7936-8196) and surface::pointGradient (src/grad.C:5697-6443), the        */
/* coupling calls of the particle complexes (src/pcomplex.C:501,            */
/* src/bd.C:1011, src/minimize.C:233).                                      */
/* ------------------------------------------------------------------------ */

#define ORC_THRESH 1e-4       /* src/interpLib.C:24 */
#define ORC_MAX_DOMAIN 100    /* src/interp.h:20 */

/* irr_kernel::domain / get_map (src/irr_kernel.C:380-393, 474-541): the
 * sub-patch level of (u, v) near the extraordinary vertex.  Returns 1 for the
 * vertex itself (u + v not > 0: the reference's domain < 0 case), -1 for a
 * level beyond MAX_DOMAIN or local coordinates outside the sub-patch (the
 * reference's exit(1) paths), else 0. */
static int irr_locate(double fu, double fv)
{
    double f_domain = -log10(fu + fv) / log10(2.0);
    if (!(f_domain < 1e300) || !(f_domain > -1e300)) return 1;
    if (fu + fv >= 1.0 - 1e-9) f_domain = 1;
    int domain = (int)lround(ceil(f_domain));
    if (domain - 1 >= ORC_MAX_DOMAIN) return -1;
    double pow2 = pow(2.0, domain - 1.0);
    fu *= pow2; fv *= pow2;
    if (fu > 0.5) { fv = 2 * fv; fu = 2.0 * fu - 1.0; }
    else if (fv > 0.5) { fu = 2 * fu; fv = 2.0 * fv - 1.0; }
    else { fv = 1 - 2 * fv; fu = 1 - 2 * fu; }
    if (fu + fv > 1.0 + ORC_THRESH || fu < -ORC_THRESH || fv < -ORC_THRESH) return -1;
    return 0;
}

/* the face's six derivative tables at (u, v) and its control points */
static int point_tables(const omesh *m, int f, double u, double v, double t[6][OMAXCP], int *ncp,
                        int *cp, double *pbc, int *irregular)
{
    int npt;
    double dudv;
    const oform *F = face_forms(m, f, &npt, &dudv);
    *ncp = F[0].ncp;
    memcpy(cp, F[0].cp, sizeof(int) * (size_t)F[0].ncp);
    memcpy(pbc, F[0].pbc, sizeof(double) * 3 * (size_t)F[0].ncp);
    *irregular = f >= m->nfr;
    if (!*irregular) {
        double B[6][12];
        orc_basis(u, v, B);
        for (int d = 0; d < 6; d++)
            for (int x = 0; x < 12; x++) t[d][x] = B[d][x];
        return 0;
    }
    int loc = irr_locate(u, v);
    if (loc) return loc;
    int val = m->val[m->face_owner[f]];
    double tab[6 * OMAXCP];
    orc_irregular_table(val, u, v, tab);
    for (int d = 0; d < 6; d++)
        for (int x = 0; x < *ncp; x++) t[d][x] = tab[d * *ncp + x];
    return 0;
}

int orc_point_eval(const omesh *m, const double *r, int f, double u, double v, double *rp,
                   double *nrm)
{
    int nv = m->nv, ncp, cp[OMAXCP], irr;
    double t[6][OMAXCP], pbc[3 * OMAXCP];
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    if (f < 0 || f >= m->nfr + m->nfi) return -1;
    int rc = point_tables(m, f, u, v, t, &ncp, cp, pbc, &irr);
    if (rc < 0) return -1;
    if (rc == 1) {
        /* the extraordinary vertex: 1/2 on it, 1/(2 val) on each neighbour at
         * its minimum image; normal unknown (src/interpLib.C:7963-8001; the
         * reference reads its vertex copy, which surface::put keeps equal to r) */
        int i = m->face_owner[f], val = m->val[i];
        double lr[3] = {0.5 * r[3 * i], 0.5 * r[3 * i + 1], 0.5 * r[3 * i + 2]};
        double w = 1.0 / (val * 2);
        for (int e = 0; e < val; e++) {
            int j = m->nb[i * OMAXV + e];
            double dr[3] = {r[3 * j] - r[3 * i], r[3 * j + 1] - r[3 * i + 1], r[3 * j + 2] - r[3 * i + 2]};
            double put[3];
            min_image(dr, (double(*)[3])m->PBC, put);
            for (int c = 0; c < 3; c++) lr[c] += (dr[c] + r[3 * i + c]) * w;
        }
        for (int c = 0; c < 3; c++) {
            rp[c] = lr[c] * al[c];
            nrm[c] = NAN;
        }
        return 0;
    }
    double X[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int x = 0; x < ncp; x++)
        for (int c = 0; c < 3; c++) {
            double q = r[3 * cp[x] + c] + pbc[3 * x + c];
            for (int d = 0; d < 3; d++) X[d][c] += t[d][x] * q;
        }
    for (int c = 0; c < 3; c++) rp[c] = X[0][c] * al[c];
    /* the normal of the unscaled tangents, as the reference forms it */
    cross3(X[1], X[2], nrm);
    normalize3(nrm);
    return 0;
}

/* pointGradient: adds to gr (3nv+3) the gradient of a point energy with
 * derivatives de_dr and de_dnrm at the surface point (f, u, v), and to
 * g_puv[2] its derivative in (u, v).  The normal's Jacobians are the
 * reference's closed forms over fac = |Ru x Rv|^3 (src/grad.C:5876-5893);
 * on_surface adds -kT dlog(g)/dr (src/grad.C:6016-6044).  On irregular faces
 * the reference forms R and its derivatives without the alpha scaling
 * (src/grad.C:6185-6220) and this restatement keeps that. */
int orc_point_grad(const omesh *m, const double *r, int f, double u, double v, const double *de_dr,
                   const double *de_dnrm, double frac_mult, int on_surface, double kT, double *gr,
                   double *g_puv)
{
    int nv = m->nv, ncp, cp[OMAXCP], irr;
    double t[6][OMAXCP], pbc[3 * OMAXCP];
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    if (f < 0 || f >= m->nfr + m->nfi) return -1;
    int rc = point_tables(m, f, u, v, t, &ncp, cp, pbc, &irr);
    if (rc != 0) return -1;   /* the reference dereferences a NULL map at the vertex */
    double X[6][3];
    memset(X, 0, sizeof X);
    for (int x = 0; x < ncp; x++)
        for (int c = 0; c < 3; c++) {
            double q = r[3 * cp[x] + c] + pbc[3 * x + c];
            if (!irr) q *= al[c];
            for (int d = 0; d < 6; d++) X[d][c] += t[d][x] * q;
        }
    const double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
    double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    double g = sqrt(RuRu * RvRv - RuRv * RuRv);
    double sq = sqrt(-pow(RuRv, 2) + RuRu * RvRv);
    double dg_RuRu = RvRv / (2. * sq), dg_RuRv = -(RuRv / sq), dg_RvRv = RuRu / (2. * sq);
    double dg_ru[3], dg_rv[3];
    for (int k = 0; k < 3; k++) {
        dg_rv[k] = dg_RvRv * 2 * Rv[k] + dg_RuRv * Ru[k];
        dg_ru[k] = dg_RuRv * Rv[k] + dg_RuRu * 2 * Ru[k];
    }
    /* J_u[j][k] = d nrm_j / d Ru_k, J_v[j][k] = d nrm_j / d Rv_k */
    double fac = pow(pow(Ru[1] * Rv[0] - Ru[0] * Rv[1], 2) + pow(Ru[2] * Rv[0] - Ru[0] * Rv[2], 2) +
                         pow(Ru[2] * Rv[1] - Ru[1] * Rv[2], 2),
                     1.5);
    double a = Ru[2] * Rv[1] - Ru[1] * Rv[2];   /* -n_x */
    double b = Ru[2] * Rv[0] - Ru[0] * Rv[2];   /*  n_y */
    double c = Ru[1] * Rv[0] - Ru[0] * Rv[1];   /* -n_z */
    double px = Ru[1] * Rv[0] * Rv[1] + Ru[2] * Rv[0] * Rv[2] - Ru[0] * (pow(Rv[1], 2) + pow(Rv[2], 2));
    double py = Rv[1] * (Ru[0] * Rv[0] + Ru[2] * Rv[2]) - Ru[1] * (pow(Rv[0], 2) + pow(Rv[2], 2));
    double pz = Ru[2] * (pow(Rv[0], 2) + pow(Rv[1], 2)) - (Ru[0] * Rv[0] + Ru[1] * Rv[1]) * Rv[2];
    double qx = pow(Ru[1], 2) * Rv[0] - Ru[0] * Ru[1] * Rv[1] + Ru[2] * (Ru[2] * Rv[0] - Ru[0] * Rv[2]);
    double qy = -(Ru[0] * Ru[1] * Rv[0]) + pow(Ru[0], 2) * Rv[1] + Ru[2] * (Ru[2] * Rv[1] - Ru[1] * Rv[2]);
    double qz = -(Ru[0] * Ru[2] * Rv[0]) + pow(Ru[0], 2) * Rv[2] + Ru[1] * (-(Ru[2] * Rv[1]) + Ru[1] * Rv[2]);
    double Ju[3][3] = {{-(a * px) / fac, b * px / fac, -(c * px) / fac},
                       {-(a * py) / fac, b * py / fac, -(c * py) / fac},
                       {a * pz / fac, -(b * pz) / fac, c * pz / fac}};
    double Jv[3][3] = {{a * qx / fac, -(b * qx) / fac, c * qx / fac},
                       {a * qy / fac, -(b * qy) / fac, c * qy / fac},
                       {a * qz / fac, -(b * qz) / fac, c * qz / fac}};
    double de_ru[3], de_rv[3], Nu[3] = {0, 0, 0}, Nv[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        de_ru[k] = (de_dnrm[0] * Ju[0][k] + de_dnrm[1] * Ju[1][k] + de_dnrm[2] * Ju[2][k]) * frac_mult;
        de_rv[k] = (de_dnrm[0] * Jv[0][k] + de_dnrm[1] * Jv[1][k] + de_dnrm[2] * Jv[2][k]) * frac_mult;
    }
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++) {
            Nu[j] += Ju[j][k] * Suu[k] + Jv[j][k] * Suv[k];
            Nv[j] += Ju[j][k] * Suv[k] + Jv[j][k] * Svv[k];
        }
    g_puv[0] += de_dr[0] * Ru[0] + de_dr[1] * Ru[1] + de_dr[2] * Ru[2];
    g_puv[1] += de_dr[0] * Rv[0] + de_dr[1] * Rv[1] + de_dr[2] * Rv[2];
    g_puv[0] += de_dnrm[0] * Nu[0] + de_dnrm[1] * Nu[1] + de_dnrm[2] * Nu[2];
    g_puv[1] += de_dnrm[0] * Nv[0] + de_dnrm[1] * Nv[1] + de_dnrm[2] * Nv[2];
    for (int x = 0; x < ncp; x++) {
        double dedr[3];
        for (int k = 0; k < 3; k++) {
            dedr[k] = de_dr[k] * frac_mult * t[0][x] + de_ru[k] * t[1][x] + de_rv[k] * t[2][x];
            if (on_surface) dedr[k] -= (1 / g) * kT * (dg_ru[k] * t[1][x] + dg_rv[k] * t[2][x]);
            gr[3 * cp[x] + k] += dedr[k] * al[k];
            gr[3 * nv + k] += dedr[k] * (r[3 * cp[x] + k] + pbc[3 * x + k]);
        }
    }
    return 0;
}

/* Point geometry: surface::ru / surface::rv (src/uv_map.C:1347-1700: the
 * alpha-scaled tangents in the face's parameters), surface::g
 * (src/uv_map.C:884-1106: |lru x lrv| of unscaled positions; on irregular
 * faces the reference scales the sub-patch derivatives by the face's own u and
 * v instead of the sub-patch factors, kept here as it computes it) and
 * surface::c (src/uv_map.C:2361-2990: principal curvatures from the shape
 * operator of the alpha-scaled surface with SAFETY, their sum and product,
 * and the principal directions in (u, v)).  out[15] = Ru[3], Rv[3], g,
 * c1 + c2, c1 c2, c1, c2, vec1[2], vec2[2].  Returns -1 where the reference
 * stops (the extraordinary vertex itself, levels beyond MAX_DOMAIN, (u, v)
 * outside an irregular face). */
int orc_point_geometry(const omesh *m, const double *r, int f, double u, double v, double *out)
{
    int nv = m->nv, ncp, cp[OMAXCP], irr;
    double t[6][OMAXCP], pbc[3 * OMAXCP];
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    if (f < 0 || f >= m->nfr + m->nfi) return -1;
    if (point_tables(m, f, u, v, t, &ncp, cp, pbc, &irr) != 0) return -1;
    double X[6][3];
    memset(X, 0, sizeof X);
    for (int x = 0; x < ncp; x++)
        for (int c = 0; c < 3; c++) {
            double q = r[3 * cp[x] + c] + pbc[3 * x + c];
            for (int d = 0; d < 6; d++) X[d][c] += t[d][x] * q;
        }
    /* g(): unscaled tangents; irregular: the sub-patch derivative (the
     * table's divided by the sub-patch factor, 2^level in magnitude) times
     * the face's u and v */
    double gu[3], gv[3], gn[3];
    double su = 1.0, sv = 1.0;
    if (irr) {
        double f_domain = -log10(u + v) / log10(2.0);
        if (u + v >= 1.0 - 1e-9) f_domain = 1;
        int domain = (int)lround(ceil(f_domain));
        double pow2 = pow(2.0, domain - 1.0);
        double fu = u * pow2, fv = v * pow2;
        double s = (fu > 0.5 || fv > 0.5) ? 2.0 : -2.0;
        su = sv = pow2 * s;
    }
    for (int c = 0; c < 3; c++) {
        gu[c] = X[1][c] / su * (irr ? u : 1.0);
        gv[c] = X[2][c] / sv * (irr ? v : 1.0);
    }
    cross3(gu, gv, gn);
    double g_ref = sqrt(gn[0] * gn[0] + gn[1] * gn[1] + gn[2] * gn[2]);
    double Ru[3], Rv[3], Suu[3], Suv[3], Svv[3], nrm[3];
    for (int c = 0; c < 3; c++) {
        Ru[c] = al[c] * X[1][c];
        Rv[c] = al[c] * X[2][c];
        Suu[c] = al[c] * X[3][c];
        Suv[c] = al[c] * X[4][c];
        Svv[c] = al[c] * X[5][c];
    }
    cross3(Ru, Rv, nrm);
    normalize3(nrm);
    double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    double g = sqrt(RuRu * RvRv - RuRv * RuRv);
    double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
    double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
    double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
    double a = 1.0 / (g * g) * (nsuu * RvRv - nsuv * RuRv);
    double b = (1.0 / (g * g)) * (nsuv * RvRv - nsvv * RuRv);
    double cc = 1.0 / (g * g) * (nsuv * RuRu - nsuu * RuRv);
    double d = (1.0 / (g * g)) * (nsvv * RuRu - nsuv * RuRv);
    double fac = sqrt(SAFETY_ + a * a + 4 * b * cc - 2 * a * d + d * d);
    double c1 = -0.5 * (a + d - fac), c2 = -0.5 * (a + d + fac);
    double v1[2], v2[2];
    if (fabs(cc) > 1e-6) {
        v1[0] = -(-a + d + fac) / (2 * cc); v1[1] = 1.0;
        v2[0] = -(-a + d - fac) / (2 * cc); v2[1] = 1.0;
    } else if (fabs(b) > 1e-6) {
        v1[0] = -(b) / (a - d); v1[1] = 1.0;
        v2[0] = 1; v2[1] = 0;
    } else {
        v1[0] = 1; v1[1] = 0;
        v2[0] = -RuRv / sqrt(RuRu) / sqrt(RvRv); v2[1] = 1;
    }
    double l1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1]), l2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
    for (int c = 0; c < 3; c++) {
        out[c] = Ru[c];
        out[3 + c] = Rv[c];
    }
    out[6] = g_ref;
    out[7] = c1 + c2;
    out[8] = c1 * c2;
    out[9] = c1;
    out[10] = c2;
    out[11] = v1[0] / l1;
    out[12] = v1[1] / l1;
    out[13] = v2[0] / l2;
    out[14] = v2[1] / l2;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Particle terms of the gradient: surface::pgrad (src/grad.C:3487-5336) and */
/* the p_face_area terms grad / igrad add on faces carrying particles       */
/* (src/grad.C:445-446, 458-459, 6819-6820)                                 */
/* ------------------------------------------------------------------------ */

/* the face's ten derivative tables at (u, v) (position, two first, three
 * second and four third derivatives, in the face's own parameters) and its
 * control points; irregular faces through the sub-patch map, derivative rows
 * scaled by the sub-patch factor as pgrad does with get_map_transform
 * (src/grad.C:4417-4418, 4614-4655) */
static int point_tables10(const omesh *m, int f, double u, double v, double t[10][OMAXCP], int *ncp, int *cp,
                          double *pbc)
{
    int npt;
    double dudv;
    const oform *F = face_forms(m, f, &npt, &dudv);
    *ncp = F[0].ncp;
    memcpy(cp, F[0].cp, sizeof(int) * (size_t)F[0].ncp);
    memcpy(pbc, F[0].pbc, sizeof(double) * 3 * (size_t)F[0].ncp);
    if (f < m->nfr) {
        double B[10][12];
        orc_basis10(u, v, B);
        for (int d = 0; d < 10; d++)
            for (int x = 0; x < 12; x++) t[d][x] = B[d][x];
        return 0;
    }
    if (irr_locate(u, v) != 0) return -1;   /* the vertex itself or outside */
    int val = m->val[m->face_owner[f]];
    double tab[10 * OMAXCP];
    if (irregular_table_nd(val, u, v, 10, tab) != 0) return -1;
    for (int d = 0; d < 10; d++)
        for (int x = 0; x < *ncp; x++) t[d][x] = tab[d * *ncp + x];
    return 0;
}

/* The backward chain of grad's energy terms, E(c1, c2, g) -> (a, b, c, d, g)
 * -> (nsuu, nsuv, nsvv, RuRu, RuRv, RvRv) -> (Ru, Rv, Suu, Suv, Svv)
 * (src/grad.C:344-494): dE[0..4] = dE/dRu, dE/dRv, dE/dSuu, dE/dSuv, dE/dSvv */
static void shape_back(const double *Ru, const double *Rv, const double *Suu, const double *Suv, const double *Svv,
                       double de_c1, double de_c2, double de_g, double dE[5][3])
{
    double nrm[3];
    cross3(Ru, Rv, nrm);
    normalize3(nrm);
    double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    double g = sqrt(RuRu * RvRv - RuRv * RuRv);
    double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
    double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
    double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
    double g2 = g * g, g3 = g * g * g;
    double a = 1.0 / g2 * (nsuu * RvRv - nsuv * RuRv);
    double b = (1.0 / g2) * (nsuv * RvRv - nsvv * RuRv);
    double c = 1.0 / g2 * (nsuv * RuRu - nsuu * RuRv);
    double d = (1.0 / g2) * (nsvv * RuRu - nsuv * RuRv);
    double S = sqrt(SAFETY_ + a * a + 4 * b * c - 2 * a * d + d * d);
    double dc1_a = -0.5 * (1 - (2 * a - 2 * d) / (2. * S));
    double dc1_b = c / S, dc1_c = b / S;
    double dc1_d = -0.5 * (1 - (-2 * a + 2 * d) / (2. * S));
    double dc2_a = -0.5 * (1 + (2 * a - 2 * d) / (2. * S));
    double dc2_b = -c / S, dc2_c = -b / S;
    double dc2_d = -0.5 * (1 + (-2 * a + 2 * d) / (2. * S));
    double de_a = de_c1 * dc1_a + de_c2 * dc2_a;
    double de_b = de_c1 * dc1_b + de_c2 * dc2_b;
    double de_c = de_c1 * dc1_c + de_c2 * dc2_c;
    double de_d = de_c1 * dc1_d + de_c2 * dc2_d;
    de_g += de_a * (-2 * (-(nsuv * RuRv) + nsuu * RvRv) / g3);
    de_g += de_b * (-2 * (-(nsvv * RuRv) + nsuv * RvRv) / g3);
    de_g += de_c * (-2 * (nsuv * RuRu - nsuu * RuRv) / g3);
    de_g += de_d * (-2 * (nsvv * RuRu - nsuv * RuRv) / g3);
    double de_nsuu = de_a * (RvRv / g2) + de_c * (-RuRv / g2);
    double de_nsuv = de_a * (-RuRv / g2) + de_b * (RvRv / g2) + de_c * (RuRu / g2) + de_d * (-RuRv / g2);
    double de_nsvv = de_b * (-RuRv / g2) + de_d * (RuRu / g2);
    double de_RuRu = de_c * (nsuv / g2) + de_d * (nsvv / g2);
    double de_RvRv = de_a * (nsuu / g2) + de_b * (nsuv / g2);
    double de_RuRv = de_a * (-nsuv / g2) + de_b * (-nsvv / g2) + de_c * (-nsuu / g2) + de_d * (-nsuv / g2);
    de_RuRu += de_g * RvRv / (2. * sqrt(-RuRv * RuRv + RuRu * RvRv));
    de_RuRv += de_g * (-(RuRv / sqrt(-RuRv * RuRv + RuRu * RvRv)));
    de_RvRv += de_g * RuRu / (2. * sqrt(-RuRv * RuRv + RuRu * RvRv));
    double de_n[3];
    for (int k = 0; k < 3; k++) {
        de_n[k] = de_nsuu * Suu[k] + de_nsuv * Suv[k] + de_nsvv * Svv[k];
        dE[0][k] = 2 * Ru[k] * de_RuRu + Rv[k] * de_RuRv;
        dE[1][k] = 2 * Rv[k] * de_RvRv + Ru[k] * de_RuRv;
        dE[2][k] = de_nsuu * nrm[k];
        dE[3][k] = de_nsuv * nrm[k];
        dE[4][k] = de_nsvv * nrm[k];
    }
    double N[3];
    cross3(Ru, Rv, N);
    double n2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
    double fac = pow(n2, 1.5);
    double J[3][3];
    for (int i2 = 0; i2 < 3; i2++)
        for (int j2 = 0; j2 < 3; j2++) J[i2][j2] = ((i2 == j2) * n2 - N[i2] * N[j2]) / fac;
    double de_N[3];
    for (int j2 = 0; j2 < 3; j2++) de_N[j2] = de_n[0] * J[0][j2] + de_n[1] * J[1][j2] + de_n[2] * J[2][j2];
    double t1[3], t2[3];
    cross3(Rv, de_N, t1);
    cross3(de_N, Ru, t2);
    for (int k = 0; k < 3; k++) {
        dE[0][k] += t1[k];
        dE[1][k] += t2[k];
    }
}

/* pgrad for one particle on face f at (u, v): E = 1/2 kc pa (c1 + c2 - pc0)^2
 * at the point; adds dE/dr (and dE/dalpha) into g and dE/du, dE/dv into pg
 * (src/grad.C:3700-3935 regular, 4614-5335 irregular) */
static int particle_pgrad(omesh *m, const double *r, int f, double u, double v, double pc0, double pa, double *g,
                          double *pg)
{
    int nv = m->nv, ncp, cp[OMAXCP];
    double t[10][OMAXCP], pbc[3 * OMAXCP];
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    if (point_tables10(m, f, u, v, t, &ncp, cp, pbc) != 0) return -1;
    double X[10][3];
    memset(X, 0, sizeof X);
    for (int x = 0; x < ncp; x++)
        for (int c = 0; c < 3; c++) {
            double q = al[c] * (r[3 * cp[x] + c] + pbc[3 * x + c]);
            for (int d = 0; d < 10; d++) X[d][c] += t[d][x] * q;
        }
    double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
    double *Suuu = X[6], *Suuv = X[7], *Suvv = X[8], *Svvv = X[9];
    double nrm[3];
    cross3(Ru, Rv, nrm);
    normalize3(nrm);
    double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    double g_ = sqrt(RuRu * RvRv - RuRv * RuRv);
    double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
    double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
    double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
    double g2 = g_ * g_, g3 = g_ * g_ * g_;
    double a = 1.0 / g2 * (nsuu * RvRv - nsuv * RuRv);
    double b = (1.0 / g2) * (nsuv * RvRv - nsvv * RuRv);
    double c = 1.0 / g2 * (nsuv * RuRu - nsuu * RuRv);
    double d = (1.0 / g2) * (nsvv * RuRu - nsuv * RuRv);
    double S = sqrt(SAFETY_ + a * a + 4 * b * c - 2 * a * d + d * d);
    double e1 = -0.5 * (a + d - S), e2 = -0.5 * (a + d + S);
    /* en = 1/2 kc pa (c1 + c2 - pc0)^2 (src/grad.C:3837-3840) */
    double de_c1 = m->prm.kc * pa * (e1 + e2 - pc0);
    double de_c2 = m->prm.kc * pa * (e1 + e2 - pc0);
    double dE[5][3];
    shape_back(Ru, Rv, Suu, Suv, Svv, de_c1, de_c2, 0.0, dE);
    for (int x = 0; x < ncp; x++)
        for (int k = 0; k < 3; k++) {
            double s = dE[0][k] * t[1][x] + dE[1][k] * t[2][x] + dE[2][k] * t[3][x] + dE[3][k] * t[4][x] +
                       dE[4][k] * t[5][x];
            g[3 * cp[x] + k] += s * al[k];
            g[3 * nv + k] += s * (r[3 * cp[x] + k] + pbc[3 * x + k]);
        }
    /* the (u, v) gradient (src/grad.C:3856-3933): every quantity's
     * derivative along u and v through the third derivatives */
    double N[3];
    cross3(Ru, Rv, N);
    double n2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
    double fac = pow(n2, 1.5);
    double Jn[3][3];   /* d nrm_i / d N_j */
    for (int i2 = 0; i2 < 3; i2++)
        for (int j2 = 0; j2 < 3; j2++) Jn[i2][j2] = ((i2 == j2) * n2 - N[i2] * N[j2]) / fac;
    /* dN/du = Suu x Rv + Ru x Suv, dN/dv = Suv x Rv + Ru x Svv */
    double a1[3], a2[3], dNu[3], dNv[3];
    cross3(Suu, Rv, a1);
    cross3(Ru, Suv, a2);
    for (int k = 0; k < 3; k++) dNu[k] = a1[k] + a2[k];
    cross3(Suv, Rv, a1);
    cross3(Ru, Svv, a2);
    for (int k = 0; k < 3; k++) dNv[k] = a1[k] + a2[k];
    double dnu[3], dnv[3];
    for (int i2 = 0; i2 < 3; i2++) {
        dnu[i2] = Jn[i2][0] * dNu[0] + Jn[i2][1] * dNu[1] + Jn[i2][2] * dNu[2];
        dnv[i2] = Jn[i2][0] * dNv[0] + Jn[i2][1] * dNv[1] + Jn[i2][2] * dNv[2];
    }
#define DOT(p, q) ((p)[0] * (q)[0] + (p)[1] * (q)[1] + (p)[2] * (q)[2])
    double d_nsuu_u = DOT(Suuu, nrm) + DOT(Suu, dnu), d_nsuu_v = DOT(Suuv, nrm) + DOT(Suu, dnv);
    double d_nsuv_u = DOT(Suuv, nrm) + DOT(Suv, dnu), d_nsuv_v = DOT(Suvv, nrm) + DOT(Suv, dnv);
    double d_nsvv_u = DOT(Suvv, nrm) + DOT(Svv, dnu), d_nsvv_v = DOT(Svvv, nrm) + DOT(Svv, dnv);
    double d_RuRu_u = 2 * DOT(Ru, Suu), d_RuRu_v = 2 * DOT(Ru, Suv);
    double d_RuRv_u = DOT(Ru, Suv) + DOT(Rv, Suu), d_RuRv_v = DOT(Ru, Svv) + DOT(Rv, Suv);
    double d_RvRv_u = 2 * DOT(Rv, Suv), d_RvRv_v = 2 * DOT(Rv, Svv);
#undef DOT
    double d_g_RuRu = RvRv / (2. * g_), d_g_RuRv = -(RuRv / g_), d_g_RvRv = RuRu / (2. * g_);
    double d_g_u = d_g_RuRu * d_RuRu_u + d_g_RuRv * d_RuRv_u + d_g_RvRv * d_RvRv_u;
    double d_g_v = d_g_RuRu * d_RuRu_v + d_g_RuRv * d_RuRv_v + d_g_RvRv * d_RvRv_v;
    /* partials of the shape operator entries (src/grad.C:364-386) */
    double a_g = -2 * (-(nsuv * RuRv) + nsuu * RvRv) / g3, a_nsuu = RvRv / g2, a_RvRv = nsuu / g2;
    double a_nsuv = -(RuRv / g2), a_RuRv = -(nsuv / g2);
    double b_g = -2 * (-(nsvv * RuRv) + nsuv * RvRv) / g3, b_nsvv = -(RuRv / g2), b_RvRv = nsuv / g2;
    double b_nsuv = RvRv / g2, b_RuRv = -(nsvv / g2);
    double c_g = -2 * (nsuv * RuRu - nsuu * RuRv) / g3, c_nsuu = -(RuRv / g2), c_RuRu = nsuv / g2;
    double c_nsuv = RuRu / g2, c_RuRv = -(nsuu / g2);
    double d_g = -2 * (nsvv * RuRu - nsuv * RuRv) / g3, d_nsvv = RuRu / g2, d_RuRu = nsvv / g2;
    double d_nsuv = -(RuRv / g2), d_RuRv = -(nsuv / g2);
    double da[2], db[2], dc[2], dd[2];
    double dnsuu[2] = {d_nsuu_u, d_nsuu_v}, dnsuv[2] = {d_nsuv_u, d_nsuv_v}, dnsvv[2] = {d_nsvv_u, d_nsvv_v};
    double dRuRu[2] = {d_RuRu_u, d_RuRu_v}, dRuRv[2] = {d_RuRv_u, d_RuRv_v}, dRvRv[2] = {d_RvRv_u, d_RvRv_v};
    double dg_[2] = {d_g_u, d_g_v};
    for (int q = 0; q < 2; q++) {
        da[q] = a_nsuu * dnsuu[q] + a_nsuv * dnsuv[q] + a_RuRv * dRuRv[q] + a_RvRv * dRvRv[q] + a_g * dg_[q];
        db[q] = b_nsvv * dnsvv[q] + b_nsuv * dnsuv[q] + b_RuRv * dRuRv[q] + b_RvRv * dRvRv[q] + b_g * dg_[q];
        dc[q] = c_nsuu * dnsuu[q] + c_nsuv * dnsuv[q] + c_RuRv * dRuRv[q] + c_RuRu * dRuRu[q] + c_g * dg_[q];
        dd[q] = d_nsvv * dnsvv[q] + d_nsuv * dnsuv[q] + d_RuRv * dRuRv[q] + d_RuRu * dRuRu[q] + d_g * dg_[q];
    }
    double dc1_a = -0.5 * (1 - (2 * a - 2 * d) / (1e-100 + 2. * S));
    double dc1_b = -(-1. * c) / S, dc1_c = -(-1. * b) / S;
    double dc1_d = -0.5 * (1 - (-2 * a + 2 * d) / (1e-100 + 2. * S));
    double dc2_a = -0.5 * (1 + (2 * a - 2 * d) / (1e-100 + 2. * S));
    double dc2_b = -(1. * c) / S, dc2_c = -(1. * b) / S;
    double dc2_d = -0.5 * (1 + (-2 * a + 2 * d) / (1e-100 + 2. * S));
    for (int q = 0; q < 2; q++) {
        double dc1 = dc1_a * da[q] + dc1_b * db[q] + dc1_c * dc[q] + dc1_d * dd[q];
        double dc2 = dc2_a * da[q] + dc2_b * db[q] + dc2_c * dc[q] + dc2_d * dd[q];
        pg[q] += de_c1 * dc1 + de_c2 * dc2;
    }
    return 0;
}

/* grad's term for a face carrying particle area pfa (particles whose c0
 * differs from the face's by 1e-7 or more): per Gauss point dE/dc1 = dE/dc2 =
 * -pfa dAf kc (c1 + c2 - c0) / face_area with dAf = g0 w and face_area = sum
 * g w of the current surface (src/grad.C:93-103, 212-224, 445-446, 458-459;
 * igrad 6477-6487, 6584, 6819-6820) */
static void face_particle_density_grad(omesh *m, int f, double pfa, const double *r, double *g)
{
    int npt, nv = m->nv;
    double dudv;
    oform *F = face_forms(m, f, &npt, &dudv);
    double al[3] = {r[3 * nv], r[3 * nv + 1], r[3 * nv + 2]};
    double face_area = 0.0;
    for (int p = 0; p < npt; p++) {
        double X[6][3];
        contract(m, F + p, r, X);
        double RuRu = X[1][0] * X[1][0] + X[1][1] * X[1][1] + X[1][2] * X[1][2];
        double RuRv = X[1][0] * X[2][0] + X[1][1] * X[2][1] + X[1][2] * X[2][2];
        double RvRv = X[2][0] * X[2][0] + X[2][1] * X[2][1] + X[2][2] * X[2][2];
        face_area += sqrt(RuRu * RvRv - RuRv * RuRv) * F[p].weight;
    }
    const double c0 = m->face_c0[f];
    for (int p = 0; p < npt; p++) {
        double X[6][3];
        contract(m, F + p, r, X);
        double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
        double nrm[3];
        cross3(Ru, Rv, nrm);
        normalize3(nrm);
        double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
        double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
        double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        double gg = sqrt(RuRu * RvRv - RuRv * RuRv);
        double nsuu = Suu[0] * nrm[0] + Suu[1] * nrm[1] + Suu[2] * nrm[2];
        double nsuv = Suv[0] * nrm[0] + Suv[1] * nrm[1] + Suv[2] * nrm[2];
        double nsvv = Svv[0] * nrm[0] + Svv[1] * nrm[1] + Svv[2] * nrm[2];
        double a = 1.0 / (gg * gg) * (nsuu * RvRv - nsuv * RuRv);
        double d = (1.0 / (gg * gg)) * (nsvv * RuRu - nsuv * RuRv);
        double sum = -(a + d);   /* e1 + e2 */
        double dAf = F[p].g0 * F[p].weight;
        double de = -pfa * dAf * m->prm.kc * (sum - c0) / face_area;
        double dE[5][3];
        shape_back(Ru, Rv, Suu, Suv, Svv, de, de, 0.0, dE);
        for (int x = 0; x < F[p].ncp; x++)
            for (int k = 0; k < 3; k++) {
                double s = dE[0][k] * F[p].t[1][x] + dE[1][k] * F[p].t[2][x] + dE[2][k] * F[p].t[3][x] +
                           dE[3][k] * F[p].t[4][x] + dE[4][k] * F[p].t[5][x];
                g[3 * F[p].cp[x] + k] += s * al[k];
                g[3 * nv + k] += s * (r[3 * F[p].cp[x] + k] + F[p].pbc[3 * x + k]);
            }
    }
}

/* the particle area each face carries in grad's density term */
static double *particle_face_area(const omesh *m, int n, const int *faces, const double *c0, const double *area)
{
    int nf = m->nfr + m->nfi;
    double *pfa = calloc((size_t)nf, sizeof(double));
    for (int i = 0; i < n; i++)
        if (fabs(c0[i] - m->face_c0[faces[i]]) >= 1e-7) pfa[faces[i]] += area[i];
    return pfa;
}

int orc_particle_grad(omesh *m, const double *r, int n, const int *faces, const double *uv, const double *c0,
                      const double *area, double *g, double *g_puv)
{
    int nf = m->nfr + m->nfi;
    for (int i = 0; i < n; i++)
        if (faces[i] < 0 || faces[i] >= nf) return -1 - i;
    double *pfa = particle_face_area(m, n, faces, c0, area);
    for (int f = 0; f < nf; f++)
        if (pfa[f] != 0.0) face_particle_density_grad(m, f, pfa[f], r, g);
    free(pfa);
    for (int i = 0; i < n; i++)
        if (particle_pgrad(m, r, faces[i], uv[2 * i], uv[2 * i + 1], c0[i], area[i], g, g_puv + 2 * i) != 0)
            return -1 - i;
    return 0;
}

/* sum over the particles of 1/2 kc pa (c1 + c2 - pc0)^2 at each particle, the
 * curvatures from orc_point_geometry: the energy whose gradient pgrad is */
double orc_particle_curvature_energy(const omesh *m, const double *r, int n, const int *faces, const double *uv,
                                     const double *c0, const double *area)
{
    double e = 0.0;
    for (int i = 0; i < n; i++) {
        double gq[15];
        if (orc_point_geometry(m, r, faces[i], uv[2 * i], uv[2 * i + 1], gq) != 0) return NAN;
        double h = gq[9] + gq[10] - c0[i];
        e += 0.5 * m->prm.kc * area[i] * h * h;
    }
    return e;
}

/* sum over the faces carrying particle area pfa of -(pfa / face_area(r_star))
 * sum_p g0 w 1/2 kc (c1 + c2 - c0)^2 at r: the function whose gradient at r =
 * r_star is grad's density term (face_area frozen at r_star) */
double orc_particle_density_frozen(omesh *m, const double *r_star, const double *r, int n, const int *faces,
                                   const double *c0, const double *area)
{
    int nf = m->nfr + m->nfi;
    double *pfa = particle_face_area(m, n, faces, c0, area);
    double e = 0.0;
    for (int f = 0; f < nf; f++) {
        if (pfa[f] == 0.0) continue;
        int npt;
        double dudv;
        oform *F = face_forms(m, f, &npt, &dudv);
        double face_area = 0.0, s = 0.0;
        for (int p = 0; p < npt; p++) {
            double X[6][3];
            contract(m, F + p, r_star, X);
            double n_[3];
            cross3(X[1], X[2], n_);
            face_area += sqrt(n_[0] * n_[0] + n_[1] * n_[1] + n_[2] * n_[2]) * F[p].weight;
        }
        for (int p = 0; p < npt; p++) {
            double X[6][3];
            contract(m, F + p, r, X);
            double *Ru = X[1], *Rv = X[2];
            double E11 = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            double E12 = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            double E22 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            double nn[3];
            cross3(Ru, Rv, nn);
            normalize3(nn);
            double L = X[3][0] * nn[0] + X[3][1] * nn[1] + X[3][2] * nn[2];
            double M = X[4][0] * nn[0] + X[4][1] * nn[1] + X[4][2] * nn[2];
            double N = X[5][0] * nn[0] + X[5][1] * nn[1] + X[5][2] * nn[2];
            /* c1 + c2 = -tr(I^-1 II) */
            double H = -(L * E22 - 2 * M * E12 + N * E11) / (E11 * E22 - E12 * E12);
            s += F[p].g0 * F[p].weight * 0.5 * m->prm.kc * (H - m->face_c0[f]) * (H - m->face_c0[f]);
        }
        e += -pfa[f] / face_area * s;
    }
    free(pfa);
    return e;
}
