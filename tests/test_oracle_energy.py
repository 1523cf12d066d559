"""Pins the oracle's energy and gradient (CPU only).

Checks that a dropped term, a wrong sign/index or a transposed operand would
fail:
* central finite differences with Richardson extrapolation (the reference's
  own check is src/grad.C:25-63) for every energy term at once (mode A with
  KA, mode B with per-face c0, kg and k_reg, the exact gradient form),
  including the alpha gradient;
* mode B's reference gradient form: its difference from the exact form is
  the finite-difference gradient of the area elements weighted by the
  curvature energy frozen at the point (src/grad.C:448, 6822-6824);
* translation invariance (sum of vertex gradients = 0) and rotation
  invariance (E unchanged, zero net torque) on a closed vesicle;
* closed forms: flat periodic sheet E = 0, g = 0; the area term vanishes at
  the setg0 state; a small-amplitude cosine on a periodic sheet gives
  1/2 kc int (lap h)^2 dA in mode B and the lipid-area weighting in mode A;
  a sphere's mode-A energy equals kc/4 H^2 sum_t w_t (Lo + Li) with the
  irregular faces' weight sum ~ 1/2;
* the per-vertex Brownian step equals r - mob * grad + amp * xi.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import meshgen, workloads


@pytest.fixture(scope="module")
def ves(mesh_dir):
    w = workloads.vesicle_642()
    s = O.OracleSurface(w.mesh_path)
    return w, s


def _fd(s, r, i, h):
    rp, rm = r.copy(), r.copy()
    rp[i] += h
    rm[i] -= h
    return (s.energy(rp) - s.energy(rm)) / (2 * h)


def _rich(s, r, i, h):
    return (4 * _fd(s, r, i, h / 2) - _fd(s, r, i, h)) / 3


def test_fd_gradient_mode_lipid(ves):
    w, s = ves
    s.set_params(**w.params)
    s.setg0(w.r0)
    g = s.grad(w.r)
    rng = np.random.default_rng(7)
    idx = list(range(0, 3 * 12 * 2)) + list(rng.integers(0, 3 * s.nv, 30)) + [3 * s.nv, 3 * s.nv + 2]
    scale = np.abs(g).max()
    for i in idx:
        h = 2e-3 if i < 3 * s.nv else 2e-6
        assert abs(_rich(s, w.r, i, h) - g[i]) < 2e-7 * scale + 1e-8 * abs(g[i]), i


def test_fd_gradient_all_terms_helfrich(mesh_dir):
    w = workloads.vesicle(2, 60.0, 0.8, seed=11, name="ves162_fd")
    s = O.OracleSurface(w.mesh_path)
    s.set_params(kc=5.0, kg=-1.3, KA=0.4, k_reg=2.0, mode=O.MODE_HELFRICH, grad_form=O.GRAD_EXACT)
    s.set_face_c0(np.random.default_rng(3).uniform(-0.02, 0.02, s.nfaces))
    s.setg0(w.r0)
    g = s.grad(w.r)
    scale = np.abs(g).max()
    for i in range(0, 3 * s.nv, 7):
        assert abs(_rich(s, w.r, i, 1e-3) - g[i]) < 2e-7 * scale, i


def _rich_frozen(s, r_star, i, h):
    def fd(hh):
        rp, rm = r_star.copy(), r_star.copy()
        rp[i] += hh
        rm[i] -= hh
        return (s.frozen_curvature_area(r_star, rp) - s.frozen_curvature_area(r_star, rm)) / (2 * hh)
    return (4 * fd(h / 2) - fd(h)) / 3


def _ref_form_cases():
    return [("ves162", lambda: workloads.vesicle(2, 60.0, 0.8, seed=11, name="ves162_fd")),
            ("defect", workloads.defect_vesicle),
            ("defect_sheet", workloads.defect_sheet)]


@pytest.mark.parametrize("case", [c[0] for c in _ref_form_cases()])
def test_reference_helfrich_gradient_term(mesh_dir, case):
    """surface::grad's Helfrich gradient adds 1/2 kc (c1+c2-c0)^2 dudv w to
    dE/dg on regular faces (src/grad.C:448) and nothing on irregular faces
    (src/grad.C:6822-6824).  So grad_reference - grad_exact must be the
    gradient in r, at r = r*, of sum_p 1/2 kc (H(r*) - c0)^2 dudv w g(r) over
    the regular faces' points with the curvature frozen at r*: checked by
    Richardson differences at every coordinate of the control points around
    an extraordinary vertex (where regular and irregular faces meet) and at a
    random sample, alpha included.  Adding the term on irregular faces, using
    g0 for g, dropping dudv or w, or freezing the wrong curvature fails it."""
    w = dict(_ref_form_cases())[case]()
    s = O.OracleSurface(w.mesh_path)
    prm = dict(kc=5.0, kg=-1.3, KA=0.4, k_reg=2.0, mode=O.MODE_HELFRICH)
    s.set_params(**prm, grad_form=O.GRAD_EXACT)
    s.set_face_c0(np.random.default_rng(3).uniform(-0.02, 0.02, s.nfaces))
    s.setg0(w.r0)
    r = w.r.copy()
    r[-3:] = [1.02, 0.99, 1.01]
    g_exact = s.grad(r)
    s.set_params(**prm, grad_form=O.GRAD_REFERENCE)
    g_ref = s.grad(r)
    diff = g_ref - g_exact
    # the term is not small: same order as the bending gradient itself
    assert np.abs(diff[:-3]).max() > 1e-3 * np.abs(g_exact[:-3]).max()
    irregular = [v for v in range(s.nv) if len(s.neighbors(v)) != 6]
    assert irregular
    idx = set()
    for v in irregular[:2]:
        for u in [v] + s.neighbors(v):
            idx.update(3 * u + k for k in range(3))
    rng = np.random.default_rng(1)
    idx.update(int(i) for i in rng.integers(0, 3 * s.nv, 24))
    idx.update([3 * s.nv, 3 * s.nv + 1, 3 * s.nv + 2])
    scale = np.abs(diff).max()
    for i in sorted(idx):
        h = 2e-3 if i < 3 * s.nv else 1e-4
        fd = _rich_frozen(s, r, i, h)
        assert abs(fd - diff[i]) < 1e-8 * scale + 1e-8 * abs(diff[i]), (case, i, fd, diff[i])


def test_grad_form_does_not_touch_lipid_mode(ves):
    """The lipid form (FIXED_A + LOCAL_LIPID_ENERGY) has one gradient."""
    w, s = ves
    s.set_params(**w.params, grad_form=O.GRAD_REFERENCE)
    s.setg0(w.r0)
    g0 = s.grad(w.r)
    s.set_params(**w.params, grad_form=O.GRAD_EXACT)
    assert np.array_equal(s.grad(w.r), g0)


@pytest.mark.parametrize("mode", [O.MODE_LIPID, O.MODE_HELFRICH])
def test_fd_gradient_valence_7(mesh_dir, mode):
    """A sphere with valence-5 and valence-7 vertices (edge flips): the
    gradient at every control point of the valence-7 faces matches Richardson
    differences of the energy, in both energy forms."""
    w = workloads.defect_vesicle()
    s = O.OracleSurface(w.mesh_path)
    if mode == O.MODE_HELFRICH:
        s.set_params(kc=5.0, kg=-1.3, KA=0.4, k_reg=2.0, mode=mode, grad_form=O.GRAD_EXACT)
        s.set_face_c0(np.random.default_rng(5).uniform(-0.02, 0.02, s.nfaces))
    else:
        s.set_params(**w.params)
    s.setg0(w.r0)
    g = s.grad(w.r)
    scale = np.abs(g).max()
    v7 = [v for v in range(s.nv) if len(s.neighbors(v)) == 7]
    assert len(v7) == 8
    idx = set()
    for v in v7[:3]:
        for u in [v] + s.neighbors(v):
            idx.update(3 * u + k for k in range(3))
    for i in sorted(idx):
        assert abs(_rich(s, w.r, i, 2e-3) - g[i]) < 2e-7 * scale + 1e-8 * abs(g[i]), i


def test_translation_rotation_invariance(ves):
    w, s = ves
    s.set_params(**w.params)
    s.setg0(w.r0)
    E = s.energy(w.r)
    g = s.grad(w.r)
    gv = g[:-3].reshape(-1, 3)
    assert np.abs(gv.sum(axis=0)).max() < 1e-12 * np.abs(gv).max() * s.nv
    X = w.r[:-3].reshape(-1, 3)
    torque = np.cross(X, gv).sum(axis=0)
    assert np.abs(torque).max() < 1e-10 * np.abs(gv).max() * 100.0 * s.nv
    th = 0.37
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    r2 = np.concatenate([(X @ Rz.T).reshape(-1), [1.0, 1.0, 1.0]])
    assert abs(s.energy(r2) - E) < 1e-11 * abs(E)
    r3 = w.r.copy()
    r3[:-3] += np.tile([3.0, -2.0, 5.0], s.nv)
    assert abs(s.energy(r3) - E) < 1e-11 * abs(E)


def test_area_term_zero_at_reference(ves):
    w, s = ves
    s.set_params(kc=0.0, KA=3.0)
    s.setg0(w.r0)
    assert abs(s.energy(w.r0)) < 1e-18
    assert s.energy(w.r) > 0


def _cos_sheet(mesh_dir, amp, nx=64):
    mesh = meshgen.hex_sheet(nx, nx, 1000.0)
    path = str(mesh_dir / ("cos_sheet_%d.mesh" % nx))
    meshgen.write_mesh(mesh, path)
    s = O.OracleSurface(path)
    r0 = meshgen.flat_coords(mesh)
    q = 2 * math.pi / 1000.0
    r = r0.copy()
    r[2:-3:3] += amp * np.cos(q * mesh.r[:, 0])
    return s, r0, r, q, mesh


def test_flat_sheet_zero(mesh_dir):
    s, r0, r, q, mesh = _cos_sheet(mesh_dir, 0.0, nx=16)
    s.set_params(kc=10.0, KA=1.0, k_reg=1.0)
    s.setg0(r0)
    assert abs(s.energy(r0)) < 1e-20
    assert np.abs(s.grad(r0)[:-3]).max() < 1e-12


def test_cosine_helfrich_closed_form(mesh_dir):
    amp, kc = 2.0, 7.0
    s, r0, r, q, mesh = _cos_sheet(mesh_dir, amp)
    s.set_params(kc=kc, mode=O.MODE_HELFRICH)
    s.setg0(r0)
    Lx, Ly = mesh.pbc[0, 0], mesh.pbc[1, 1]
    exact = 0.5 * kc * amp ** 2 * q ** 4 * Lx * Ly / 2
    assert abs(s.energy(r) / exact - 1) < 0.02
    # mode A weights the curvature by the lipid areas instead of the metric
    s.set_params(kc=kc, mode=O.MODE_LIPID)
    s.setg0(r0)
    At = Lx * Ly / s.nfaces
    Lo, Li = s.tri_lipids(0)
    assert Lo == Li == 68.3 * max(1, math.floor((At - 32.5) / 68.3) + 1)
    assert abs(s.energy(r) / (exact * (Lo + Li) / (2 * At)) - 1) < 0.02


def test_sphere_lipid_energy_weights(mesh_dir):
    R, kc = 400.0, 3.0
    w = workloads.vesicle(4, R, 0.0, seed=0, name="sphere_l4")
    s = O.OracleSurface(w.mesh_path)
    s.set_params(kc=kc)
    s.setg0(w.r0)
    c = kc / 4 * (2 / R) ** 2
    reg = sum(sum(s.tri_lipids(t)) for t in range(s.ntri()) if s.tri_face(t) < s.nreg)
    irr = sum(sum(s.tri_lipids(t)) for t in range(s.ntri()) if s.tri_face(t) >= s.nreg)
    e_reg = sum(s.face_energy(f, w.r0) for f in range(s.nreg))
    e_irr = sum(s.face_energy(f, w.r0) for f in range(s.nreg, s.nfaces))
    # regular faces: H = 2/R on the (nearly spherical) limit surface, weights
    # sum to 1 and the lipid term carries no dudv
    assert abs(e_reg / (c * reg) - 1) < 0.005
    # irregular faces: their 16 weights sum to ~1/2 (the reference's quirk),
    # so they carry roughly half a regular face's curvature energy
    assert 0.55 < e_irr / (c * irr) < 0.75
    assert abs(s.energy(w.r0) - (e_reg + e_irr)) < 1e-12 * abs(e_reg)


def test_bd_step_composition(ves):
    w, s = ves
    s.set_params(**w.params)
    s.setg0(w.r0)
    rng = np.random.default_rng(5)
    xi = rng.standard_normal((2, 3 * s.nv))
    r1, en = s.bd_steps(w.r, xi, 0.01, 0.3, want_energies=True)
    r = w.r.copy()
    for k in range(2):
        assert en[k] == s.energy(r)
        g = s.grad(r)
        r[:-3] = r[:-3] + (-0.01 * g[:-3] + 0.3 * xi[k])
    np.testing.assert_array_equal(r1, r)
