"""GPU: particle energies on the limit surface (surface::penergy,
src/interpLib.C:8330-8790, with the face densities of fenergy / ifenergy,
3612-3613 and 3850-3851) against the oracle, through the C-ABI.

Bars, FP64, relative to the largest well-conditioned term of the batch:
1e-12 on regular faces; on irregular faces at sub-patch level d (u + v >=
0.05) 1e-12 times 4^d, the conditioning of second derivatives formed from the
level's map entries times 4^d over sums that cancel (measured: 8e-14 on
regular faces, 2e-13, 8e-13, 2e-12, 2e-11 at levels 1-4); nearer the
extraordinary vertex 1000 times the oracle's own change for a one-ulp move
of u or v.  Covered: both energy
forms, random face c0, the vesicle and the periodic sheet with valence-5 and
-7 vertices, particles at a face's Gauss points cancelling its density, the
batch split over host threads, and rejections.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu

REG_PTS = [(0.1, 0.1), (0.3, 0.1), (0.5, 0.1), (0.7, 0.1), (0.9, 0.1), (0.1, 0.3), (0.3, 0.3),
           (0.5, 0.3), (0.7, 0.3), (0.1, 0.5), (0.3, 0.5), (0.5, 0.5), (0.1, 0.7), (0.3, 0.7),
           (0.1, 0.9)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(w, mode, seed=3):
    prm = dict(w.params)
    prm.update(kc=14.0, kg=2.0, mode=mode)
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**prm)
    c0 = np.random.default_rng(seed).uniform(-0.02, 0.02, o.nfaces)
    o.set_face_c0(c0)
    o.setg0(w.r0)
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **prm)
    g.generatePlan()
    g.set_face_c0(c0)
    g.setg0(w.r0)
    return g, o, c0


def _particles(o, n, seed, deep=0):
    rng = np.random.default_rng(seed)
    faces = rng.integers(0, o.nfaces, n).astype(np.int32)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    uv = np.stack([u, v], 1)
    irr = np.flatnonzero(faces >= o.nreg)
    for k, i in enumerate(irr[:deep]):
        s = 10.0 ** -(1 + k % 8)
        uv[i] *= s / max(uv[i].sum(), 1e-300)
    c0 = rng.uniform(-0.03, 0.03, n)
    area = rng.uniform(20.0, 80.0, n)
    return faces, uv, c0, area


def _check(g, o, r, faces, uv, c0, area):
    E = g.particle_energy(r, faces, uv, c0, area)
    Eo = o.particle_energy(r, faces, uv, c0, area)
    ok = (faces < o.nreg) | (uv.sum(axis=1) >= 0.05)
    scale = np.abs(Eo[ok]).max()
    err = np.abs(E - Eo)
    reg = faces < o.nreg
    t = uv.sum(axis=1)
    level = np.where(reg, 0, np.where(t >= 1 - 1e-9, 1, np.ceil(-np.log2(np.maximum(t, 1e-300)))))
    say("   n=%d scale %.3g; max err regular %.3g, irregular by level %s" % (
        len(faces), scale, err[reg].max(),
        {int(l): float("%.3g" % err[ok & ~reg & (level == l)].max()) for l in np.unique(level[ok & ~reg])}))
    assert (err[ok] <= 1e-12 * scale * 4.0 ** level[ok]).all()
    if (~ok).any():
        base = Eo[~ok]
        sens = np.zeros_like(base)
        for k in range(2):
            q = uv[~ok].copy()
            q[:, k] = np.nextafter(q[:, k], np.inf)
            sens = np.maximum(sens, np.abs(o.particle_energy(r, faces[~ok], q, c0[~ok], area[~ok]) - base))
        say("   near the vertex: max err %.3g, max 1-ulp response %.3g" % (err[~ok].max(), sens.max()))
        assert (err[~ok] <= 1e-12 * scale + 1000 * sens).all()
    return E


@pytest.mark.parametrize("mode", [O.MODE_LIPID, O.MODE_HELFRICH])
def test_particle_energy_matches_oracle(mode):
    w = workloads.vesicle_642()
    g, o, _ = _pair(w, mode)
    faces, uv, c0, area = _particles(o, 3000, 5, deep=60)
    say("[particles] vesicle mode=%d" % mode)
    _check(g, o, w.r, faces, uv, c0, area)


def test_particle_energy_periodic_defect_sheet():
    w = workloads.defect_sheet()
    g, o, _ = _pair(w, O.MODE_HELFRICH)
    faces, uv, c0, area = _particles(o, 2000, 6)
    say("[particles] defect sheet")
    _check(g, o, w.r, faces, uv, c0, area)


@pytest.mark.parametrize("mode", [O.MODE_LIPID, O.MODE_HELFRICH])
def test_gauss_point_particles_cancel(mode):
    """Particles at a regular face's Gauss points with the points' areas
    dudv g0 w and the face's c0 take out what the density puts in."""
    w = workloads.vesicle_642()
    g, o, c0 = _pair(w, mode)
    say("[particles] Gauss-point cancellation mode=%d" % mode)
    for f in [0, 11, o.nreg // 2, o.nreg - 1]:
        pa = np.array([0.5 * o.face_table(f, p)[3] * o.face_table(f, p)[1] for p in range(15)])
        out = g.particle_energy(w.r, np.full(15, f, np.int32), np.array(REG_PTS), np.full(15, c0[f]), pa)
        num, _ = o.face_density(w.r, f)
        assert abs(out.sum()) <= 1e-12 * abs(num)


def test_threaded_batch_and_rejections():
    w = workloads.vesicle_642()
    g, o, _ = _pair(w, O.MODE_LIPID)
    faces, uv, c0, area = _particles(o, 100000, 7)
    say("[particles] threaded batch n=%d" % len(faces))
    E = g.particle_energy(w.r, faces, uv, c0, area)
    for k in range(0, len(faces), 25000):
        s = slice(k, k + 25000)
        assert np.array_equal(g.particle_energy(w.r, faces[s], uv[s], c0[s], area[s]), E[s])
    Eo = o.particle_energy(w.r, faces[:500], uv[:500], c0[:500], area[:500])
    assert np.abs(E[:500] - Eo).max() <= 1e-12 * np.abs(Eo).max()
    bad = faces.copy()
    bad[60000] = o.nfaces
    with pytest.raises(S.UmcpError, match="point 60000: face index out of range"):
        g.particle_energy(w.r, bad, uv, c0, area)
    # the surface still answers afterwards
    assert np.array_equal(g.particle_energy(w.r, faces[:10], uv[:10], c0[:10], area[:10]), E[:10])
    # the densities need the reference state (setg0)
    h = S.Surface(w.mesh_path)
    h.set_params(precision=S.PREC_FP64, **w.params)
    h.generatePlan()
    with pytest.raises(S.UmcpError, match="setg0"):
        h.particle_energy(w.r, faces[:10], uv[:10], c0[:10], area[:10])
    h.close()
