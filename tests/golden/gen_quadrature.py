"""Writes the irregular-face quadrature fixture (tests/golden/quadrature.json).

The reference builds the 4 x 4 Gauss points of an irregular face with
Golub-Welsch over two singular weights (src/gauss.C:96-299), integrating the
recurrence moments with GSL's ``gsl_integration_qags(f, 0, hi, 1e-9, 0, 100)``
(src/gauss.C:132-135, 234-237) and diagonalising the Jacobi matrix with LAPACK
``dsyev`` (src/gauss.C:178, 280).  GSL and LAPACK are not vendored in the
reference (configure.ac:65-67, 74-150; versions unpinned) and are absent here,
so this script uses their published algorithms as shipped in scipy:
``scipy.integrate.quad`` is QUADPACK's dqagse (the routine GSL's qags
transcribes) with the same absolute tolerance 1e-9, relative 0 and limit 100,
and ``numpy.linalg.eigh`` is LAPACK's symmetric eigensolver.  It then applies
the reference's point and weight transforms (src/interpLib.C:5610-5645).

An mpmath evaluation of the same moments at 40 digits is stored beside it so
tests can see how far QAGS at 1e-9 sits from the exact integrals.

Run:  python tests/golden/gen_quadrature.py  (writes quadrature.json here)
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
from scipy.integrate import quad

KVALS = {4: 0.375, 5: 4.52254248593737e-01, 7: 5.30872450464684e-01,
         8: 5.51776695296637e-01, 9: 5.66511110779745e-01, 10: 5.77254248593736e-01}


def _weight_v(k):
    c = math.log(2.0) / (math.log(32.0) + 4.0 * math.log(k))
    return lambda v: -(-1 + k ** (4.0 * math.log2(v)) * v ** 5) * c


def _weight_u(k, v):
    def f(u):
        val = (1.0 / (u + v)) * k ** (-math.log2(u + v))
        return 1.0 / (val * val * val * val)
    return f


def golub_welsch(weight, hi, n=4):
    a = np.zeros(n + 1)
    b = np.zeros(n + 1)
    prev, mu0 = 1.0, 0.0
    for i in range(n + 1):
        def poly(x, i=i):
            pm1, p = 0.0, 1.0
            for j in range(i):
                p, pm1 = (x - a[j]) * p - b[j] * pm1, p
            return p
        num = quad(lambda x: x * poly(x) ** 2 * weight(x), 0.0, hi,
                   epsabs=1e-9, epsrel=0.0, limit=100)[0]
        den = quad(lambda x: poly(x) ** 2 * weight(x), 0.0, hi,
                   epsabs=1e-9, epsrel=0.0, limit=100)[0]
        a[i] = num / den
        b[i] = den / prev if i > 0 else 0.0
        if i == 0:
            mu0 = den
        prev = den
    J = np.diag(a[:n]) + np.diag(np.sqrt(b[1:n]), 1) + np.diag(np.sqrt(b[1:n]), -1)
    x, V = np.linalg.eigh(J)
    w = mu0 * V[0, :] ** 2 / np.sum(V * V, axis=0)
    return x, w


def irregular_points(val):
    k = KVALS[val]
    vx, vw = golub_welsch(_weight_v(k), 1.0)
    pts = []
    u_nodes = []
    for a in range(4):
        fv = vx[a]
        v_factor = -(-1 + math.pow(k, 4.0 * math.log2(fv)) * fv ** 5) * math.log(2.0) / \
            (math.log(32.0) + 4.0 * math.log(k))
        v_weight = vw[a] / v_factor
        ux, uw = golub_welsch(_weight_u(k, fv), 1.0 - fv)
        u_nodes.append((ux.tolist(), uw.tolist()))
        for b in range(4):
            fu = ux[b]
            factor = math.pow(1.0 / (fu + fv) * math.pow(k, -math.log2(fu + fv)), 4.0)
            pts.append((fu, fv, v_weight * (uw[b] * factor)))
    return dict(k=k, v_nodes=vx.tolist(), v_weights=vw.tolist(), u=u_nodes,
                points=[list(p) for p in pts])


def main():
    out = {"source": "scipy.integrate.quad (QUADPACK dqagse, epsabs 1e-9, epsrel 0, "
                     "limit 100) + numpy.linalg.eigh",
           "valences": {}}
    for val in sorted(KVALS):
        out["valences"][str(val)] = irregular_points(val)
    try:
        import mpmath as mp
        mp.mp.dps = 40
        exact = {}
        for val in (5, 7):
            k = mp.mpf(KVALS[val])
            c = mp.log(2) / (mp.log(32) + 4 * mp.log(k))
            wv = lambda v: -(-1 + k ** (4 * mp.log(v, 2)) * v ** 5) * c  # noqa: E731
            moments = [float(mp.quad(lambda v: v ** m * wv(v), [0, 1])) for m in range(9)]
            exact[str(val)] = moments
        out["exact_v_moments"] = exact
    except ImportError:
        pass
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quadrature.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
