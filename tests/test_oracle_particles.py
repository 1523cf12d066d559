"""The oracle's particle energy (surface::penergy, src/interpLib.C:8330-8790,
with the face densities of fenergy / ifenergy, 3850-3851 and 3612-3613),
pinned against quantities other tests fix:

- with kg = k_reg = KA = 0 in the Helfrich form the density's numerator of
  every face is the face's energy (faceEnergy, pinned against the
  reference's golden values and finite differences), and its area the face's
  reference area (dudv g0 w summed, as surface::area's A0);
- the regular faces' density does not depend on the energy form (kc 1/2
  (c1 + c2 - c0)^2 in both), while the irregular faces' follows the point
  energy density: kg c1 c2 alone in the lipid form, so with kc = 0 and
  kg = 1 the irregular numerator is the face energy and the regular one is 0;
- particles at a regular face's 15 Gauss points carrying the points' areas
  dudv g0 w and the face's c0 have zero total energy (the face density
  removes exactly what they add);
- the particle term is linear in the particle's area, zero at zero area, and
  a quadratic in its c0 with minimum at c1 + c2 of the point.
"""
import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import workloads

REG_PTS = [(0.1, 0.1), (0.3, 0.1), (0.5, 0.1), (0.7, 0.1), (0.9, 0.1), (0.1, 0.3), (0.3, 0.3),
           (0.5, 0.3), (0.7, 0.3), (0.1, 0.5), (0.3, 0.5), (0.5, 0.5), (0.1, 0.7), (0.3, 0.7),
           (0.1, 0.9)]


@pytest.fixture(scope="module")
def ves():
    w = workloads.vesicle_642()
    o = O.OracleSurface(w.mesh_path)
    o.setg0(w.r0)
    c0 = np.random.default_rng(3).uniform(-0.02, 0.02, o.nfaces)
    o.set_face_c0(c0)
    return w, o, c0


def _faces(o):
    return [0, 7, o.nreg // 2, o.nreg - 1, o.nreg, o.nreg + 17, o.nreg + 33, o.nfaces - 1]


def test_density_is_the_helfrich_face_energy(ves):
    w, o, _ = ves
    o.set_params(kc=14.0, kg=0.0, KA=0.0, k_reg=0.0, mode=O.MODE_HELFRICH)
    _, a0_tot = o.area(w.r)
    s_area = 0.0
    for f in range(o.nfaces):
        num, area = o.face_density(w.r, f)
        e = o.face_energy(f, w.r)
        assert abs(num - e) <= 1e-13 * abs(e)
        s_area += area
    assert abs(s_area - a0_tot) <= 1e-12 * a0_tot


def test_density_regular_form_irregular_point_energy(ves):
    w, o, _ = ves
    # regular faces: the same numerator in both forms
    o.set_params(kc=14.0, kg=3.0, KA=0.5, k_reg=0.1, mode=O.MODE_LIPID)
    lip = [o.face_density(w.r, f) for f in _faces(o)]
    o.set_params(kc=14.0, kg=0.0, KA=0.0, k_reg=0.0, mode=O.MODE_HELFRICH)
    for f, (num, _) in zip(_faces(o), lip):
        if f < o.nreg:
            e = o.face_energy(f, w.r)
            assert abs(num - e) <= 1e-13 * abs(e)
    # irregular faces in the lipid form: kg c1 c2 only
    o.set_params(kc=0.0, kg=1.0, KA=0.0, k_reg=0.0, mode=O.MODE_LIPID)
    for f in _faces(o):
        num, _ = o.face_density(w.r, f)
        if f < o.nreg:
            assert num == 0.0
        else:
            e = o.face_energy(f, w.r)
            assert abs(num - e) <= 1e-13 * abs(e) and abs(e) > 0


@pytest.mark.parametrize("mode", [O.MODE_LIPID, O.MODE_HELFRICH])
def test_gauss_point_particles_cancel(ves, mode):
    w, o, c0 = ves
    o.set_params(kc=14.0, kg=2.0, KA=0.0, k_reg=0.0, mode=mode)
    for f in [0, 7, o.nreg // 2, o.nreg - 1]:
        num, area = o.face_density(w.r, f)
        pa = np.array([0.5 * o.face_table(f, p)[3] * o.face_table(f, p)[1] for p in range(15)])
        assert abs(pa.sum() - area) <= 1e-14 * area
        out = o.particle_energy(w.r, [f] * 15, REG_PTS, [c0[f]] * 15, pa)
        assert abs(out.sum()) <= 1e-12 * abs(num)
        assert abs(num) > 0


def test_particle_term_shape(ves):
    w, o, _ = ves
    o.set_params(kc=14.0, kg=2.0, KA=0.0, k_reg=0.0, mode=O.MODE_LIPID)
    f, uv = o.nreg + 5, (0.3, 0.2)
    H = o.point_geometry(w.r, [f], [uv])[0, 7]
    num, area = o.face_density(w.r, f)
    e0 = o.particle_energy(w.r, [f], [uv], [H], [0.0])[0]
    assert e0 == 0.0
    a1 = o.particle_energy(w.r, [f], [uv], [0.01], [1.0])[0]
    a3 = o.particle_energy(w.r, [f], [uv], [0.01], [3.0])[0]
    assert abs(a3 - 3 * a1) <= 1e-13 * abs(a3)
    # quadratic in c0, minimum at the point's c1 + c2, value -a rho_f there
    at = o.particle_energy(w.r, [f], [uv], [H], [2.0])[0]
    assert num != 0.0 and abs(at + 2.0 * num / area) <= 1e-13 * abs(at)
    for d in (1e-3, 1e-2):
        ep = o.particle_energy(w.r, [f], [uv], [H + d], [2.0])[0]
        em = o.particle_energy(w.r, [f], [uv], [H - d], [2.0])[0]
        q = 0.5 * 14.0 * 2.0 * d * d
        assert abs((ep - at) - q) <= 1e-6 * q
        assert abs(ep - em) <= 1e-6 * q
