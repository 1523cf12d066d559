"""Pins the oracle's lipid compositions (CPU only).

* its restatement of glibc rand() equals libc's rand() after srand(seed) for
  seeds across the 32-bit range (the reference packs lipids with rand(),
  src/interpLib.C:3127-3196, 3286-3358);
* a one-type composition packs and weighs exactly as the default single type;
* packing invariants: every leaflet holds more than A - 65/2 of lipid area,
  and dropping one lipid of some type present leaves at most that much (the
  loop stops at the first draw past the target); each type's area is a whole
  number of its lipids; over the mesh the types' area fractions follow the
  input parts within sampling noise;
* energy: two types identical in c0 and area per lipid give the single-type
  energy; on a sphere the lipid energy of the regular faces is
  kc/4 sum_t sum_x ((H - c0_x)^2 Lo_tx + (H + c0_x)^2 Li_tx) with H = 2/R,
  which fixes the leaflet signs; finite differences pin the gradient.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import workloads

MIX = dict(c0=[-1.0 / 87.3, -1.0 / 29.4, -1.0 / 27.2], apl=[68.3, 63.4, 31.0],
           inner=[3.0, 2.0, 1.0], outer=[5.0, 0.0, 2.0])


@pytest.mark.parametrize("seed", [1, 7, 12345, 2 ** 31 - 1, 0, 2 ** 31, 3000000000, 2 ** 32 - 1])
def test_glibc_rand_restatement(seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand.argtypes = [ctypes.c_uint]
    libc.srand(seed)
    ref = np.array([libc.rand() for _ in range(2000)])
    assert np.array_equal(O.glibc_rand(seed, 2000), ref)


@pytest.fixture(scope="module")
def ves():
    return workloads.vesicle_642()


def _pair(w, comp=None, seed=3, **prm):
    s = O.OracleSurface(w.mesh_path)
    s.set_params(**{**w.params, **prm})
    if comp:
        s.set_lipids(comp["c0"], comp["apl"], comp["inner"], comp["outer"], seed)
    s.setg0(w.r0)
    return s


def test_one_type_composition_is_the_default(ves):
    a = _pair(ves, lipid_c0=0.012)
    b = _pair(ves, dict(c0=[0.012], apl=[68.3], inner=[1.0], outer=[2.0]), lipid_c0=0.012)
    for t in range(a.ntri()):
        assert a.tri_lipids(t) == b.tri_lipids(t)
    assert a.energy(ves.r) == b.energy(ves.r)
    assert np.array_equal(a.grad(ves.r), b.grad(ves.r))


def test_packing_invariants(ves):
    s = _pair(ves, MIX)
    apl = np.array(MIX["apl"])
    tot_o = np.zeros(3)
    tot_i = np.zeros(3)
    for t in range(s.ntri()):
        f = s.tri_face(t)
        info = s.face_info(f)
        A = sum(s.face_table(f, p)[3] * s.face_table(f, p)[1] * (0.5 if f < s.nreg else 1.0)
                for p in range(info["npt"]))
        lo, li = s.tri_lipid_types(t)
        for L in (lo, li):
            n = L / apl
            assert np.allclose(n, np.round(n), atol=1e-9) and (n >= 0).all()
            assert L.sum() > A - 32.5
            present = L > 0
            # the lipid added last (one of the types present) crossed the target
            assert (L.sum() - apl[present]).min() <= max(A - 32.5, 0.0) + 1e-9
        if f < s.nreg:
            # no outer parts of the second type: one draw per round picks by
            # the cumulative parts.  Irregular faces draw afresh for every type
            # (src/interpLib.C:3338-3341), so there it can be picked.
            assert lo[1] == 0.0
        tot_o += lo
        tot_i += li
    # area fractions follow the number fractions weighted by area per lipid
    for tot, parts in ((tot_o, MIX["outer"]), (tot_i, MIX["inner"])):
        expect = np.array(parts) * apl / np.dot(parts, apl)
        assert np.abs(tot / tot.sum() - expect).max() < 0.05


def test_identical_types_reduce_to_one(ves):
    one = _pair(ves, lipid_c0=-0.02, lipid_apl=60.0)
    two = _pair(ves, dict(c0=[-0.02, -0.02], apl=[60.0, 60.0], inner=[1.0, 1.0], outer=[1.0, 3.0]),
                lipid_c0=-0.02, lipid_apl=60.0)
    for t in range(one.ntri()):
        assert np.isclose(sum(one.tri_lipids(t)), sum(two.tri_lipids(t)), rtol=0, atol=1e-9)
    E1, E2 = one.energy(ves.r), two.energy(ves.r)
    assert abs(E1 - E2) <= 1e-13 * abs(E1)


def test_sphere_multi_lipid_energy(mesh_dir):
    R, kc = 400.0, 3.0
    w = workloads.vesicle(4, R, 0.0, seed=0, name="sphere_l4")
    comp = dict(c0=[0.01, -0.008], apl=[50.0, 70.0], inner=[1.0, 2.0], outer=[3.0, 1.0])
    s = _pair(w, comp, seed=5, kc=kc)
    H = 2.0 / R
    c0 = np.array(comp["c0"])
    pred = 0.0
    for t in range(s.ntri()):
        if s.tri_face(t) < s.nreg:
            lo, li = s.tri_lipid_types(t)
            pred += kc / 4 * (np.dot((H - c0) ** 2, lo) + np.dot((H + c0) ** 2, li))
    e_reg = sum(s.face_energy(f, w.r0) for f in range(s.nreg))
    assert abs(e_reg / pred - 1) < 0.005
    # the opposite leaflet convention would be far off
    wrong = sum(kc / 4 * (np.dot((H + c0) ** 2, s.tri_lipid_types(t)[0]) +
                          np.dot((H - c0) ** 2, s.tri_lipid_types(t)[1]))
                for t in range(s.ntri()) if s.tri_face(t) < s.nreg)
    assert abs(e_reg / wrong - 1) > 0.05


def _rich(s, r, i, h):
    def fd(hh):
        rp, rm = r.copy(), r.copy()
        rp[i] += hh
        rm[i] -= hh
        return (s.energy(rp) - s.energy(rm)) / (2 * hh)
    return (4 * fd(h / 2) - fd(h)) / 3


def test_fd_gradient_multi_lipid(ves):
    s = _pair(ves, MIX, KA=0.3)
    g = s.grad(ves.r)
    rng = np.random.default_rng(3)
    idx = list(range(0, 36)) + list(rng.integers(0, 3 * s.nv, 24))
    scale = np.abs(g).max()
    for i in idx:
        assert abs(_rich(s, ves.r, i, 2e-3) - g[i]) < 2e-7 * scale + 1e-8 * abs(g[i]), i
