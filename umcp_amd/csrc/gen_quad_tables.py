"""Generates quad_tables.inc: the 16 Gauss points of an irregular face.

For each valence the reference integrates over the irregular face with a
4 x 4 product rule whose nodes come from Golub-Welsch over two singular weights
(src/gauss.C:27-299), the moments integrated by GSL's QAGS at absolute
tolerance 1e-9 (src/gauss.C:132-135), then mapped to (u, v, weight) by
src/interpLib.C:5610-5645.  These points depend only on the valence, so the
library ships them as constants computed here with QUADPACK's dqagse (the
routine GSL's qags transcribes) via scipy, at the same tolerances, and LAPACK's
symmetric eigensolver via numpy.

Run:  python umcp_amd/csrc/gen_quad_tables.py   (rewrites quad_tables.inc)
"""
import math
import os

import numpy as np
from scipy.integrate import quad

KVALS = {4: 0.375, 5: 4.52254248593737e-01, 7: 5.30872450464684e-01,
         8: 5.51776695296637e-01, 9: 5.66511110779745e-01, 10: 5.77254248593736e-01}


def moments_rule(weight, hi, n=4):
    """Recurrence of the monic orthogonal polynomials, then the Jacobi matrix."""
    alpha = np.zeros(n + 1)
    beta = np.zeros(n + 1)
    norm_prev, mu0 = 1.0, 0.0
    for i in range(n + 1):
        def p_i(x, i=i):
            p_prev, p = 0.0, 1.0
            for j in range(i):
                p, p_prev = (x - alpha[j]) * p - beta[j] * p_prev, p
            return p
        kw = dict(epsabs=1e-9, epsrel=0.0, limit=100)
        first = quad(lambda x: x * p_i(x) ** 2 * weight(x), 0.0, hi, **kw)[0]
        norm = quad(lambda x: p_i(x) ** 2 * weight(x), 0.0, hi, **kw)[0]
        alpha[i] = first / norm
        beta[i] = norm / norm_prev if i else 0.0
        if i == 0:
            mu0 = norm
        norm_prev = norm
    jac = np.diag(alpha[:n])
    off = np.sqrt(beta[1:n])
    jac += np.diag(off, 1) + np.diag(off, -1)
    nodes, vecs = np.linalg.eigh(jac)
    return nodes, mu0 * vecs[0, :] ** 2 / np.sum(vecs * vecs, axis=0)


def points_for(val):
    k = KVALS[val]
    c = math.log(2.0) / (math.log(32.0) + 4.0 * math.log(k))

    def wv(v):
        return -(-1 + k ** (4.0 * math.log2(v)) * v ** 5) * c
    vn, vw = moments_rule(wv, 1.0)
    pts = []
    for a in range(4):
        fv = vn[a]
        vweight = vw[a] / wv(fv)

        def wu(u, fv=fv):
            s = u + fv
            return (s * k ** math.log2(s)) ** 4
        un, uw = moments_rule(wu, 1.0 - fv)
        for b in range(4):
            fu = un[b]
            factor = math.pow(1.0 / (fu + fv) * math.pow(k, -math.log2(fu + fv)), 4.0)
            pts.append((fu, fv, vweight * (uw[b] * factor)))
    return pts


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    lines = ["// Generated by gen_quad_tables.py: irregular-face Gauss points per valence,",
             "// {u, v, weight}, 16 per valence (4 v nodes outer, 4 u nodes inner).",
             "static const int kQuadValences[] = {%s};" % ", ".join(str(v) for v in sorted(KVALS)),
             "static const double kQuadPoints[][16][3] = {"]
    for val in sorted(KVALS):
        lines.append("  { // valence %d" % val)
        for u, v, w in points_for(val):
            lines.append("    {%s, %s, %s}," % (repr(float(u)), repr(float(v)), repr(float(w))))
        lines.append("  },")
    lines.append("};")
    with open(os.path.join(here, "quad_tables.inc"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
