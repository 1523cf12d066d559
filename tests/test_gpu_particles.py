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

The particle terms of the gradient (umcp_particle_grad: surface::pgrad and
grad's particle-area term) against the oracle's restatement, pinned there by
finite differences: vertex and alpha gradient norm-wise 1e-12 relative with
particles on regular faces only; with particles on irregular faces at
sub-patch level d 1e-12 times 8^d (the third derivatives' conditioning;
measured 1.3e-12 at levels up to 3), the (u, v) gradient per particle 1e-12
times 8^d of the largest; bitwise repeatable, accumulating, and the
extraordinary vertex rejected.
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


# --- particle terms of the gradient (surface::pgrad, src/grad.C:3487-5336, and
# grad's particle-area term, src/grad.C:445-446, 458-459, 6819-6820) ---------

def _grad_check(g, o, r, faces, uv, c0, area, tag):
    gg, gp = g.particle_grad(r, faces, uv, c0, area)
    og, op = o.particle_grad(r, faces, uv, c0, area)
    rel = np.linalg.norm(gg[:-3] - og[:-3]) / np.linalg.norm(og[:-3])
    rel_a = np.abs(gg[-3:] - og[-3:]).max() / np.abs(og[-3:]).max()
    gp, op = gp.reshape(-1, 2), op.reshape(-1, 2)
    t = uv.sum(axis=1)
    reg = faces < o.nreg
    level = np.where(reg, 0, np.where(t >= 1 - 1e-9, 1, np.ceil(-np.log2(np.maximum(t, 1e-300)))))
    scale = np.abs(op).max()
    err = np.abs(gp - op).max(axis=1)
    say("   %s: n=%d vertex grad rel %.3g alpha rel %.3g, (u,v) max err %.3g of %.3g (levels %s)" % (
        tag, len(faces), rel, rel_a, err.max(), scale, sorted(set(level.astype(int).tolist()))))
    # third derivatives of a level-d sub-patch are formed from map entries
    # times 8^d over sums that cancel
    bar = 1e-12 * 8.0 ** level.max()
    assert rel <= bar and rel_a <= bar
    assert (err <= 1e-12 * scale * 8.0 ** level).all()
    return gg, gp


def _grad_particles(o, n, seed):
    """random particles; irregular ones at sub-patch levels 1-4 (u + v >= 1/16)"""
    faces, uv, c0, area = _particles(o, n, seed)
    irr = faces >= o.nreg
    t = uv[irr].sum(axis=1, keepdims=True)
    uv[irr] *= np.maximum(t, 1.0 / 15.0) / np.maximum(t, 1e-300)
    return faces, uv, c0, area


@pytest.mark.parametrize("mode", [O.MODE_LIPID, O.MODE_HELFRICH])
def test_particle_grad_vesicle(mode):
    w = workloads.vesicle_642()
    g, o, _ = _pair(w, mode)
    faces, uv, c0, area = _grad_particles(o, 600, 21)
    irr = np.flatnonzero(faces >= o.nreg)
    assert irr.size >= 10
    r = w.r.copy()
    r[-3:] = [1.01, 0.99, 1.02]
    say("[particle grad] vesicle 642 mode", mode)
    _grad_check(g, o, r, faces, uv, c0, area, "random c0 (particle-area term on)")
    reg = faces < o.nreg
    _grad_check(g, o, r, faces[reg], uv[reg], c0[reg], area[reg], "regular faces only (bar 1e-12)")
    # particles whose c0 equals their face's: pgrad alone
    fc0 = np.random.default_rng(3).uniform(-0.02, 0.02, o.nfaces)
    _grad_check(g, o, r, faces, uv, fc0[faces], area, "face c0 (pgrad alone)")


def test_particle_grad_defect_sheet():
    """valence-5 and -7 faces reaching across the periodic box"""
    w = workloads.defect_sheet()
    g, o, _ = _pair(w, O.MODE_HELFRICH)
    faces, uv, c0, area = _grad_particles(o, 500, 5)
    say("[particle grad] defect sheet")
    _grad_check(g, o, w.r, faces, uv, c0, area, "defect sheet")


def test_particle_grad_accumulates_repeats_and_rejects():
    w = workloads.vesicle_642()
    g, o, _ = _pair(w, O.MODE_HELFRICH)
    faces, uv, c0, area = _grad_particles(o, 300, 9)
    g1, p1 = g.particle_grad(w.r, faces, uv, c0, area)
    g2, p2 = g.particle_grad(w.r, faces, uv, c0, area)
    assert np.array_equal(g1, g2) and np.array_equal(p1, p2)
    acc, accp = np.ones_like(g1), np.full(p1.shape, 2.0)
    g.particle_grad(w.r, faces, uv, c0, area, acc, accp)
    assert np.array_equal(acc, 1.0 + g1) and np.array_equal(accp, 2.0 + p1)
    # the extraordinary vertex itself has no sub-patch (the reference stops)
    irr = int(np.flatnonzero(faces >= o.nreg)[0])
    bad_uv = uv.copy()
    bad_uv[irr] = (0.0, 0.0)
    with pytest.raises(S.UmcpError, match="point %d" % irr):
        g.particle_grad(w.r, faces, bad_uv, c0, area)
    # the handle still works and gives the same result
    g3, p3 = g.particle_grad(w.r, faces, uv, c0, area)
    assert np.array_equal(g3, g1) and np.array_equal(p3, p1)
