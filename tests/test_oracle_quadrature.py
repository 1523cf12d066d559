"""Pins the oracle's restatement of the irregular-face quadrature.

The reference takes its 4 x 4 irregular Gauss points from Golub-Welsch over
QUADPACK/GSL QAGS moments (src/gauss.C:96-299).  Checks:
* the oracle's C QAGS + Jacobi nodes equal scipy's QUADPACK dqagse + LAPACK
  eigh (tests/golden/quadrature.json) to ~1e-14;
* QAGS at epsabs 1e-9 sits within 1e-9 of the exact (mpmath) moments;
* the known weight sums of the 16 irregular points (0.501039 at valence 5,
  0.499741 at valence 7) that make irregular faces carry half weight;
* Gauss quadrature exactness: the 4 v-nodes integrate polynomials of degree
  <= 7 against the v weight exactly (closed form of the moments).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "quadrature.json")))


@pytest.mark.parametrize("val", ["4", "5", "7", "8", "9", "10"])
def test_nodes_match_quadpack(val):
    d = GOLD["valences"][val]
    x, w = O.v_points(d["k"])
    np.testing.assert_allclose(x, d["v_nodes"], rtol=0, atol=5e-14)
    np.testing.assert_allclose(w, d["v_weights"], rtol=0, atol=5e-14)
    for a in range(4):
        ux, uw = O.u_points(d["k"], d["v_nodes"][a])
        np.testing.assert_allclose(ux, d["u"][a][0], rtol=0, atol=5e-14)
        np.testing.assert_allclose(uw, d["u"][a][1], rtol=0, atol=5e-14)


def test_weight_sums_half():
    s5 = sum(p[2] for p in GOLD["valences"]["5"]["points"])
    s7 = sum(p[2] for p in GOLD["valences"]["7"]["points"])
    assert abs(s5 - 0.501039) < 5e-7
    assert abs(s7 - 0.499741) < 5e-7


@pytest.mark.parametrize("val", ["5", "7"])
def test_gauss_exactness_closed_form(val):
    """v weight = (1 - v^q)/q with q = 5 + 4 log2 k: moments in closed form."""
    k = GOLD["valences"][val]["k"]
    q = 5.0 + 4.0 * math.log2(k)
    x, w = O.v_points(k)
    for m in range(8):
        exact = (1.0 / (m + 1) - 1.0 / (m + q + 1)) / q
        assert abs(np.dot(w, x ** m) - exact) < 2e-9
    # and the mpmath moments stored with the fixture agree with the closed form
    for m, mom in enumerate(GOLD["exact_v_moments"][val]):
        exact = (1.0 / (m + 1) - 1.0 / (m + q + 1)) / q
        assert abs(mom - exact) < 1e-14


def test_qags_known_integrals():
    import ctypes as C
    L = O.lib()
    F = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
    L.orc_qags.argtypes = [F, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                           C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_qags.restype = C.c_double
    err, ier, last = C.c_double(), C.c_int(), C.c_int()
    # endpoint singularity: int_0^1 log(x)/sqrt(x) = -4 (needs extrapolation)
    f = F(lambda x, _: math.log(x) / math.sqrt(x))
    v = L.orc_qags(f, None, 0.0, 1.0, 0.0, 1e-10, 100, C.byref(err), C.byref(ier), C.byref(last))
    assert abs(v + 4.0) < 1e-9 and ier.value == 0 and last.value > 1
    g = F(lambda x, _: math.exp(x))
    v = L.orc_qags(g, None, 0.0, 1.0, 1e-12, 0.0, 100, C.byref(err), C.byref(ier), C.byref(last))
    assert abs(v - (math.e - 1)) < 1e-14
