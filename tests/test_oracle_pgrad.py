"""The oracle's particle gradient: surface::pgrad (src/grad.C:3487-5336) and the
particle-area terms grad and igrad add on faces carrying particles
(src/grad.C:445-446, 458-459, 6819-6820), pinned by finite differences of
independent energies:

- pgrad is the gradient of sum_i 1/2 kc a_i (c1 + c2 - c0_i)^2 with the
  curvatures at the particles taken from the point geometry (a separate
  forward path): Richardson differences in every control-point coordinate of
  the faces carrying particles, in alpha, and in each particle's (u, v), on a
  vesicle (regular faces and irregular faces at sub-patch levels 1-5) and on
  the periodic sheet with valence-5 and -7 vertices;
- grad's particle-area term is the gradient, at r = r*, of
  -sum_f (P_f / A_f(r*)) sum_p g0 w 1/2 kc (c1 + c2 - c0_f)^2 over the faces
  whose particles' c0 differ from the face's (A_f = sum g w of the current
  surface, frozen at r*); particles whose c0 equals the face's add no such
  term (src/grad.C:100-101).
A dropped third-derivative term, the sub-patch scale, or the density term's
frozen area would each fail one of these.
"""
import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import workloads


def _particles(o, rng, n, irregular_levels=(1, 2, 3, 5)):
    """n particles: regular faces at random (u, v), irregular faces at points
    of the given sub-patch levels (u + v in (2^-L, 2^(1-L)))."""
    faces, uv = [], []
    nreg = o.nfaces - o.nirr
    for i in range(n):
        if o.nirr and i % 3 == 2:
            f = nreg + int(rng.integers(0, o.nirr))
            L = irregular_levels[i % len(irregular_levels)]
            s = rng.uniform(2.0 ** -L * 1.05, 2.0 ** (1 - L) * 0.95)
            t = rng.uniform(0.15, 0.85)
            faces.append(f)
            uv.append((s * t, s * (1 - t)))
        else:
            f = int(rng.integers(0, nreg))
            a, b = rng.uniform(0.05, 0.9, 2)
            if a + b > 0.95:
                a, b = 0.95 - b * 0.5, b * 0.5
            faces.append(f)
            uv.append((a, b))
    return np.array(faces, dtype=np.int32), np.array(uv)


def _rich(fn, x, i, h):
    def fd(hh):
        xp, xm = x.copy(), x.copy()
        xp[i] += hh
        xm[i] -= hh
        return (fn(xp) - fn(xm)) / (2 * hh)
    return (4 * fd(h / 2) - fd(h)) / 3


def _cases():
    return {"vesicle": lambda: workloads.vesicle(2, 60.0, 0.8, seed=11, name="ves162_fd"),
            "defect_sheet": workloads.defect_sheet}


@pytest.mark.parametrize("case", ["vesicle", "defect_sheet"])
def test_pgrad_is_gradient_of_particle_curvature_energy(mesh_dir, case):
    w = _cases()[case]()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(kc=5.0, mode=O.MODE_HELFRICH)
    o.setg0(w.r0)
    rng = np.random.default_rng(4)
    faces, uv = _particles(o, rng, 12)
    area = rng.uniform(20.0, 80.0, faces.size)
    # particle c0 equal to the face's: no particle-area term, pgrad alone
    face_c0 = rng.uniform(-0.02, 0.02, o.nfaces)
    o.set_face_c0(face_c0)
    c0 = face_c0[faces]
    r = w.r.copy()
    r[-3:] = [1.01, 0.98, 1.02]
    g, gp = o.particle_grad(r, faces, uv, c0, area)
    scale = np.abs(g).max()
    assert scale > 0
    idx = set()
    for f in faces:
        for x in o.face_info(int(f))["cp"]:
            idx.update(3 * x + k for k in range(3))
    idx = sorted(idx)[:150] + [3 * o.nv, 3 * o.nv + 1, 3 * o.nv + 2]
    energy = lambda rr: o.particle_curvature_energy(rr, faces, uv, c0, area)  # noqa: E731
    for i in idx:
        h = 1e-3 if i < 3 * o.nv else 1e-4
        fd = _rich(energy, r, i, h)
        assert abs(fd - g[i]) < 1e-7 * scale + 1e-7 * abs(g[i]), (case, i, fd, g[i])
    # (u, v): differences of the particle's own energy along its parameters
    guv = gp.reshape(-1, 2)
    for p in range(faces.size):
        for q in range(2):
            def e_uv(x, p=p):
                u2 = uv.copy()
                u2[p] = x
                return o.particle_curvature_energy(r, faces, u2, c0, area)
            h = 1e-5 * max(1.0, 1.0 / (uv[p].sum()))
            fd = _rich(e_uv, uv[p].copy(), q, min(h, 0.2 * uv[p].sum()))
            assert abs(fd - guv[p, q]) < 1e-6 * np.abs(guv).max() + 1e-6 * abs(guv[p, q]), (case, p, q, fd, guv[p, q])


@pytest.mark.parametrize("case", ["vesicle", "defect_sheet"])
def test_particle_area_term_of_grad(mesh_dir, case):
    w = _cases()[case]()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(kc=5.0, mode=O.MODE_HELFRICH)
    o.setg0(w.r0)
    rng = np.random.default_rng(8)
    faces, uv = _particles(o, rng, 9)
    area = rng.uniform(20.0, 80.0, faces.size)
    c0p = rng.uniform(-0.05, 0.05, faces.size)
    face_c0 = rng.uniform(-0.02, 0.02, o.nfaces)
    r = w.r.copy()
    r[-3:] = [1.01, 0.98, 1.02]
    # pgrad alone: the faces' c0 set to their particles' (one particle per face here)
    assert len(set(faces.tolist())) == faces.size
    fc_eq = face_c0.copy()
    fc_eq[faces] = c0p
    o.set_face_c0(fc_eq)
    g_p, gp_p = o.particle_grad(r, faces, uv, c0p, area)
    o.set_face_c0(face_c0)
    g_t, gp_t = o.particle_grad(r, faces, uv, c0p, area)
    assert np.array_equal(gp_p, gp_t)   # the area term has no (u, v) part
    dens = g_t - g_p
    assert np.abs(dens).max() > 1e-3 * np.abs(g_p).max()
    frozen = lambda rr: o.particle_density_frozen(r, rr, faces, c0p, area)  # noqa: E731
    idx = set()
    for f in faces:
        for x in o.face_info(int(f))["cp"]:
            idx.update(3 * x + k for k in range(3))
    idx = sorted(idx)[:120] + [3 * o.nv, 3 * o.nv + 1, 3 * o.nv + 2]
    scale = np.abs(dens).max()
    for i in idx:
        h = 1e-3 if i < 3 * o.nv else 1e-4
        fd = _rich(frozen, r, i, h)
        assert abs(fd - dens[i]) < 1e-7 * scale + 1e-7 * abs(dens[i]), (case, i, fd, dens[i])
