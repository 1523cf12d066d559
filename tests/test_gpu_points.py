"""GPU: points on the limit surface (surface::evaluateRNRM,
src/interpLib.C:7936-8196; surface::pointGradient, src/grad.C:5697-6443)
against the oracle.

The device forms the normal's Jacobian as a projector applied to cross
products, the oracle with the reference's closed forms over |Ru x Rv|^3, so
the two agree to rounding only.  Bars, FP64 with per-vertex sums in point
order on both sides: positions to 1e-12 of the surface's extent everywhere;
normals to 1e-12 absolute and the (u, v), vertex and alpha gradients to 1e-12
of their largest entry wherever the evaluation is well conditioned (regular
faces; irregular faces at u + v >= 0.05).  Closer to an extraordinary vertex
the reference's sub-patch evaluation (weights of level d scaled by 2^d over
entries that cancel to (2 lambda)^d) is ill conditioned: one ulp of u moves
the oracle's own normal by ~5e-10 at u + v = 1e-5 and by ~0.1 at 1e-12.  There
the bar is 1000 times the oracle's own response to a one-ulp change of u or v.
Covered: regular faces, sub-patch levels 1-40 on irregular faces, the vertex
itself (evaluation only), alpha != 1, frac_mult, the on_surface term, a
periodic sheet (image shifts), accumulation into the caller's arrays,
repeatability, and the reference's rejections; the control-point weights of
get_pt_coeffs / get_pt_dcoeffs (src/uv_map.C:14-420) against the oracle's
position and tangents and against pointGradient's coupling.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import meshgen, workloads  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(w):
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    return g, o


@pytest.fixture(scope="module")
def ves():
    w = workloads.vesicle_642()
    g, o = _pair(w)
    return w, g, o


def random_points(o, n, seed, deep=0):
    """n points uniform over the faces; the first `deep` points on irregular
    faces moved towards the extraordinary vertex, u + v = 10^-1 .. 10^-12."""
    rng = np.random.default_rng(seed)
    faces = rng.integers(0, o.nfaces, n)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    uv = np.stack([u, v], 1)
    irr = np.flatnonzero(faces >= o.nreg)
    for k, i in enumerate(irr[:deep]):
        s = 10.0 ** -(1 + k % 12)
        uv[i] *= s / max(uv[i].sum(), 1e-300)
    return faces.astype(np.int32), uv


def well_conditioned(o, faces, uv):
    return (faces < o.nreg) | (uv.sum(axis=1) >= 0.05)


def ulp_response(fn, uv):
    """Largest change of fn(uv) when u or v of every point moves by one ulp."""
    base = fn(uv)
    out = np.zeros_like(base)
    for k in range(2):
        q = uv.copy()
        q[:, k] = np.nextafter(q[:, k], np.inf)
        out = np.maximum(out, np.abs(fn(q) - base))
    return out


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("alpha", [(1.0, 1.0, 1.0), (1.1, 0.9, 1.05)])
def test_eval_matches_oracle(ves, alpha):
    w, g, o = ves
    r = w.r.copy()
    r[-3:] = alpha
    faces, uv = random_points(o, 3000, 5, deep=120)
    ok = well_conditioned(o, faces, uv)
    say("[points] eval n=%d (%d near a vertex) alpha=%s" % (len(faces), (~ok).sum(), alpha))
    R, N = g.point_eval(r, faces, uv)
    Ro, No = o.point_eval(r, faces, uv)
    say("   R rel", rel(R, Ro), "nrm abs (well conditioned)", np.abs(N - No)[ok].max())
    assert rel(R, Ro) <= 1e-12
    assert np.abs(N - No)[ok].max() <= 1e-12
    sens = ulp_response(lambda q: o.point_eval(r, faces[~ok], q)[1], uv[~ok])
    err = np.abs(N[~ok] - No[~ok])
    say("   near the vertex: max err", err.max(), "max 1-ulp response", sens.max())
    assert (err <= 1e-12 + 1000 * sens).all()


def test_eval_at_extraordinary_vertex(ves):
    w, g, o = ves
    r = w.r.copy()
    faces = np.array([o.nreg, o.nreg + 5, o.nfaces - 1], np.int32)
    uv = np.zeros((3, 2))
    R, N = g.point_eval(r, faces, uv)
    Ro, No = o.point_eval(r, faces, uv)
    assert rel(R, Ro) <= 1e-12
    assert np.isnan(N).all() and np.isnan(No).all()


@pytest.mark.parametrize("alpha,on_surface", [((1.0, 1.0, 1.0), False), ((1.0, 1.0, 1.0), True),
                                              ((1.1, 0.9, 1.05), True)])
def test_grad_matches_oracle(ves, alpha, on_surface):
    w, g, o = ves
    r = w.r.copy()
    r[-3:] = alpha
    faces, uv = random_points(o, 2000, 6)
    rng = np.random.default_rng(7)
    a, b = rng.normal(size=(len(faces), 3)), rng.normal(size=(len(faces), 3))
    fm = rng.random(len(faces)) + 0.5
    kT = 0.616
    ok = well_conditioned(o, faces, uv)
    faces, uv, a, b, fm = faces[ok], uv[ok], a[ok], b[ok], fm[ok]
    say("[points] grad n=%d alpha=%s on_surface=%d" % (len(faces), alpha, on_surface))
    G, P = g.point_grad(r, faces, uv, a, b, frac_mult=fm, on_surface=on_surface, kT=kT)
    Go, Po = o.point_grad(r, faces, uv, a, b, frac_mult=fm, on_surface=int(on_surface), kT=kT)
    say("   g rel", rel(G, Go), "alpha rel", rel(G[-3:], Go[-3:]), "g_puv rel", rel(P, Po))
    assert rel(G[:-3], Go[:-3]) <= 1e-12
    assert rel(G[-3:], Go[-3:]) <= 1e-12
    assert rel(P, Po) <= 1e-12


def test_grad_near_extraordinary_vertex(ves):
    w, g, o = ves
    r = w.r.copy()
    faces, uv = random_points(o, 4000, 15, deep=10 ** 6)
    keep = ~well_conditioned(o, faces, uv)
    faces, uv = faces[keep][:240], uv[keep][:240]
    n = len(faces)
    rng = np.random.default_rng(16)
    a, b = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    say("[points] grad near the vertex n=%d, u+v down to %.1e" % (n, uv.sum(axis=1).min()))
    G, P = g.point_grad(r, faces, uv, a, b, on_surface=True, kT=0.6)
    Go, Po = o.point_grad(r, faces, uv, a, b, on_surface=1, kT=0.6)
    sG = ulp_response(lambda q: o.point_grad(r, faces, q, a, b, on_surface=1, kT=0.6)[0], uv)
    sP = ulp_response(lambda q: o.point_grad(r, faces, q, a, b, on_surface=1, kT=0.6)[1], uv)
    say("   g err", np.abs(G - Go).max(), "1-ulp response", sG.max(), "g_puv err", np.abs(P - Po).max(),
        "1-ulp response", sP.max())
    assert (np.abs(G - Go) <= 1e-12 * np.abs(Go).max() + 1000 * sG.max()).all()
    assert (np.abs(P - Po) <= 1e-12 * np.abs(Po).max() + 1000 * sP).all()


def test_grad_accumulates_and_repeats(ves):
    w, g, o = ves
    r = w.r.copy()
    faces, uv = random_points(o, 500, 8, deep=40)
    rng = np.random.default_rng(9)
    a, b = rng.normal(size=(500, 3)), rng.normal(size=(500, 3))
    G1, P1 = g.point_grad(r, faces, uv, a, b)
    G2, P2 = g.point_grad(r, faces, uv, a, b)
    assert np.array_equal(G1, G2) and np.array_equal(P1, P2)
    base = rng.normal(size=r.size)
    pb = rng.normal(size=(500, 2))
    G3, P3 = g.point_grad(r, faces, uv, a, b, g=base.copy(), g_puv=pb.copy())
    assert np.allclose(G3, base + G1, rtol=0, atol=1e-13 * np.abs(G1).max())
    assert np.allclose(P3, pb + P1, rtol=0, atol=1e-13 * np.abs(P1).max())
    # two half batches accumulated give the whole batch
    Gh, Ph = g.point_grad(r, faces[:250], uv[:250], a[:250], b[:250])
    g.point_grad(r, faces[250:], uv[250:], a[250:], b[250:], g=Gh)
    assert rel(Gh, G1) <= 1e-12


def test_periodic_sheet_matches_oracle():
    w = workloads.sheet(16, seed=3)
    g, o = _pair(w)
    r = w.r.copy()
    faces, uv = random_points(o, 1500, 10)
    say("[points] periodic sheet nv=%d faces=%d" % (o.nv, o.nfaces))
    R, N = g.point_eval(r, faces, uv)
    Ro, No = o.point_eval(r, faces, uv)
    assert rel(R, Ro) <= 1e-12 and np.abs(N - No).max() <= 1e-12
    rng = np.random.default_rng(12)
    a, b = rng.normal(size=(1500, 3)), rng.normal(size=(1500, 3))
    G, P = g.point_grad(r, faces, uv, a, b, on_surface=True, kT=0.6)
    Go, Po = o.point_grad(r, faces, uv, a, b, on_surface=1, kT=0.6)
    say("   g rel", rel(G, Go), "g_puv rel", rel(P, Po))
    assert rel(G[:-3], Go[:-3]) <= 1e-12 and rel(G[-3:], Go[-3:]) <= 1e-12 and rel(P, Po) <= 1e-12


@pytest.mark.parametrize("mesh", ["sphere", "sheet"])
def test_valence_7_faces_match_oracle(mesh):
    """Points on a sphere with valence-5 and valence-7 vertices: the
    valence-7 sub-patch maps (13 control points) against the oracle; and on
    a periodic sheet whose irregular faces reach across the box."""
    w = workloads.defect_vesicle() if mesh == "sphere" else workloads.defect_sheet()
    g, o = _pair(w)
    r = w.r.copy()
    faces, uv = random_points(o, 3000, 17, deep=60)
    ok = well_conditioned(o, faces, uv)
    say("[points] valence 5/7 %s: %d irregular points" % (mesh, int((faces >= o.nreg).sum())))
    R, N = g.point_eval(r, faces, uv)
    Ro, No = o.point_eval(r, faces, uv)
    assert rel(R, Ro) <= 1e-12 and np.abs(N - No)[ok].max() <= 1e-12
    # every irregular face at its extraordinary vertex
    irr = np.arange(o.nreg, o.nfaces, dtype=np.int32)
    R0, _ = g.point_eval(r, irr, np.zeros((len(irr), 2)))
    R0o, _ = o.point_eval(r, irr, np.zeros((len(irr), 2)))
    assert rel(R0, R0o) <= 1e-12
    rng = np.random.default_rng(18)
    f2, uv2 = faces[ok], uv[ok]
    a, b = rng.normal(size=(len(f2), 3)), rng.normal(size=(len(f2), 3))
    G, P = g.point_grad(r, f2, uv2, a, b, on_surface=True, kT=0.6)
    Go, Po = o.point_grad(r, f2, uv2, a, b, on_surface=1, kT=0.6)
    say("   g rel", rel(G, Go), "g_puv rel", rel(P, Po))
    assert rel(G[:-3], Go[:-3]) <= 1e-12 and rel(G[-3:], Go[-3:]) <= 1e-12 and rel(P, Po) <= 1e-12


def test_large_batch_is_sum_of_parts(ves):
    w, g, o = ves
    r = w.r.copy()
    n = 200000
    faces, uv = random_points(o, n, 13)
    ok = well_conditioned(o, faces[:2000], uv[:2000])
    rng = np.random.default_rng(14)
    a, b = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    say("[points] batch n=%d" % n)
    R, N = g.point_eval(r, faces, uv)
    Ro, No = o.point_eval(r, faces[:2000], uv[:2000])
    assert rel(R[:2000], Ro) <= 1e-12 and np.abs(N[:2000] - No)[ok].max() <= 1e-12
    G, P = g.point_grad(r, faces, uv, a, b)
    acc = np.zeros_like(r)
    for k in range(0, n, 50000):
        g.point_grad(r, faces[k:k + 50000], uv[k:k + 50000], a[k:k + 50000], b[k:k + 50000], g=acc)
    say("   whole vs parts rel", rel(G, acc))
    assert rel(G, acc) <= 1e-12
    _, Po = o.point_grad(r, faces[:2000], uv[:2000], a[:2000], b[:2000])
    assert rel(P[:2000][ok], Po[ok]) <= 1e-12


def test_rejections(ves):
    w, g, o = ves
    r = w.r.copy()
    z = np.zeros((1, 3))
    irr = np.array([o.nreg], np.int32)
    with pytest.raises(S.UmcpError):
        g.point_grad(r, irr, np.zeros((1, 2)), z, z)
    with pytest.raises(S.UmcpError):
        g.point_eval(r, irr, np.array([[1e-40, 1e-40]]))
    with pytest.raises(S.UmcpError):
        g.point_eval(r, np.array([o.nfaces], np.int32), np.array([[0.2, 0.2]]))
    with pytest.raises(S.UmcpError):
        g.point_eval(r, irr, np.array([[0.9, 0.3]]))
    # the surface still answers afterwards
    R, _ = g.point_eval(r, irr, np.array([[0.2, 0.2]]))
    Ro, _ = o.point_eval(r, irr, np.array([[0.2, 0.2]]))
    assert rel(R, Ro) <= 1e-12


@pytest.mark.parametrize("mesh,alpha", [("vesicle", (1.0, 1.0, 1.0)), ("vesicle", (1.1, 0.9, 1.05)),
                                        ("defect_sphere", (1.0, 1.0, 1.0)), ("defect_sheet", (1.0, 1.0, 1.0))])
def test_geometry_matches_oracle(ves, mesh, alpha):
    """surface::ru, rv, g and c at random well-conditioned points.  Bars:
    tangents, the reference's area element and c1 + c2 to 1e-12 of their
    scale s.  The individual principal curvatures are
    -(a + d)/2 -+ sqrt(SAFETY + (a - d)^2 + 4bc)/2: a rounding change dD of the
    discriminant moves them by dD / (4 |c1 - c2|), which grows without bound at
    umbilic points (a sphere is umbilic everywhere), and the principal
    directions are undefined there.  So c1, c2 and c1 c2 are held to
    1e-12 s + 1e-14 s^2 / |c1 - c2| and the unit directions to
    1e-10 + 1e-14 s / |c1 - c2|, per point.  On an irregular face at
    sub-patch level d the second-derivative weights are the level's map
    entries times 4^d over sums that cancel, so the two implementations'
    rounding of the maps reaches the curvatures multiplied by 4^d (measured:
    regular faces 1e-13, levels 2-4 up to 6e-12): every curvature bar carries
    that factor."""
    if mesh == "vesicle":
        w, g, o = ves
    else:
        w = workloads.defect_vesicle() if mesh == "defect_sphere" else workloads.defect_sheet()
        g, o = _pair(w)
    r = w.r.copy()
    r[-3:] = alpha
    faces, uv = random_points(o, 2000, 31)
    ok = well_conditioned(o, faces, uv)
    faces, uv = faces[ok], uv[ok]
    say("[points] geometry %s alpha=%s n=%d" % (mesh, alpha, len(faces)))
    G = g.point_geometry(r, faces, uv)
    Go = o.point_geometry(r, faces, uv)
    tang = np.abs(G[:, :6] - Go[:, :6]).max() / np.abs(Go[:, :6]).max()
    area = np.abs(G[:, 6] - Go[:, 6]).max() / np.abs(Go[:, 6]).max()
    s = np.abs(Go[:, 9:11]).max()
    t = uv.sum(axis=1)
    level = np.where(faces >= o.nreg, np.where(t >= 1 - 1e-9, 1, np.ceil(-np.log2(t))), 0)
    cond = 4.0 ** level
    dH = np.abs(G[:, 7] - Go[:, 7])
    split = np.maximum(np.abs(Go[:, 9] - Go[:, 10]), 1e-300)
    dc = np.abs(G[:, 9:11] - Go[:, 9:11]).max(axis=1)
    dk = np.abs(G[:, 8] - Go[:, 8])
    dv = np.abs(G[:, 11:] - Go[:, 11:]).max(axis=1)
    reg = faces < o.nreg
    say("   tangents", tang, "g", area, "c1+c2 regular", dH[reg].max() / s, "irregular",
        (dH[~reg].max() / s) if (~reg).any() else 0.0, "min |c1-c2|/s", split.min() / s,
        "directions max", dv.max())
    assert tang <= 1e-12 and area <= 1e-12
    assert (dH <= 1e-12 * s * cond).all()
    assert (dc <= (1e-12 * s + 1e-14 * s * s / split) * cond).all()
    assert (dk <= (1e-12 * s * s + 1e-14 * s ** 3 / split) * cond).all()
    assert (dv <= 1e-10 + 1e-14 * s / split * cond).all()


def test_geometry_rejects_the_vertex(ves):
    w, g, o = ves
    with pytest.raises(S.UmcpError):
        g.point_geometry(w.r, np.array([o.nreg], np.int32), np.zeros((1, 2)))


@pytest.mark.parametrize("mesh", ["vesicle", "defect_sheet"])
def test_next_face_walk(ves, mesh):
    """surface::nextFace (src/uv_map.C:1695-1961), one step per call: a move
    that stays inside the face just adds; a move that leaves it stops at the
    border it reaches first, where the new face evaluates to the same surface
    point (the reference's MOVE_CHECK condition, here to 1e-10 A), and the
    rest (du', dv') satisfies the reference's two equations
    (du r_u + dv r_v) . r_u = (du' r_u' + dv' r_v') . r_u (and with r_v) with
    the oracle's tangents; the momentum goes through the same matrix; a walk
    repeated until no displacement remains ends on a face."""
    if mesh == "vesicle":
        w, g, o = ves
    else:
        w = workloads.defect_sheet()
        g, o = _pair(w)
    r = w.r.copy()
    rng = np.random.default_rng(51)
    n = 600
    faces, uv = random_points(o, n, 52)
    duv = rng.normal(0.0, 0.4, (n, 2))
    duv[:50] *= 1e-3                                    # these stay inside
    mom = duv.copy()
    f1, uv1, d1, m1, _ = g.next_face(r, faces, uv, duv, mom=mom)
    inside = (uv + duv).min(axis=1) > 0
    inside &= (uv + duv).sum(axis=1) < 1
    assert np.array_equal(f1[inside], faces[inside])
    assert np.allclose(uv1[inside], (uv + duv)[inside], rtol=0, atol=1e-15) and not d1[inside].any()
    moved = ~inside
    say("[points] next_face %s: %d crossings" % (mesh, int(moved.sum())))
    # the crossing on the old face: the smallest positive t to a border
    u, v = uv[moved, 0], uv[moved, 1]
    du, dv = duv[moved, 0], duv[moved, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.stack([np.where(dv < 0, -v / dv, np.inf), np.where(du + dv > 0, (1 - u - v) / (du + dv), np.inf),
                      np.where(du < 0, -u / du, np.inf)], 1).min(axis=1)
    uc, vc = u + t * du, v + t * dv
    border = np.minimum(np.minimum(uc, vc), 1 - uc - vc)
    assert np.abs(border).max() < 1e-12
    Ro, _ = o.point_eval(r, faces[moved], np.stack([uc, vc], 1))
    Rn, _ = o.point_eval(r, f1[moved], uv1[moved])
    dR = Ro - Rn
    if mesh == "defect_sheet":   # the same point of a periodic surface, at its minimum image
        L = np.array(meshgen.read_mesh(w.mesh_path).pbc).diagonal()
        dR -= L * np.round(dR / L)
    say("   continuity max |dR|", np.abs(dR).max())
    assert np.abs(dR).max() < 1e-10
    assert (f1[moved] != faces[moved]).all()
    Go = o.point_geometry(r, faces[moved], np.stack([uc, vc], 1))
    Gn = o.point_geometry(r, f1[moved], uv1[moved])
    rem = np.stack([du * (1 - t), dv * (1 - t)], 1)
    lhs = rem[:, :1] * Go[:, 0:3] + rem[:, 1:] * Go[:, 3:6]
    rhs = d1[moved, :1] * Gn[:, 0:3] + d1[moved, 1:] * Gn[:, 3:6]
    for k in (0, 3):
        a = (lhs * Go[:, k:k + 3]).sum(axis=1)
        b = (rhs * Go[:, k:k + 3]).sum(axis=1)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(lhs).max() * np.abs(Go[:, k:k + 3]).max()
    # mom was set to the displacement: it comes out as the new displacement
    # times 1/(1 - t) (the map is linear and took the remaining part)
    assert np.allclose(m1[moved] * (1 - t)[:, None], d1[moved], rtol=1e-9, atol=1e-12)
    # walk to the end
    f, q, d = f1, uv1, d1
    for _ in range(50):
        if not d.any():
            break
        f, q, d, _, _ = g.next_face(r, f, q, d)
    assert not d.any()
    assert (q.min(axis=1) >= -1e-12).all() and (q.sum(axis=1) <= 1 + 1e-12).all()


def _coupled(r, ids, w, d):
    """alpha * sum_x w[:, d, x] r[ids[:, x]] (no image shifts: closed surfaces)."""
    nv = (r.size - 3) // 3
    X = r[:3 * nv].reshape(nv, 3)
    m = ids >= 0
    out = np.einsum("nx,nxk->nk", np.where(m, w[:, d], 0.0), X[np.where(m, ids, 0)])
    return out * r[-3:]


@pytest.mark.parametrize("alpha", [(1.0, 1.0, 1.0), (1.1, 0.9, 1.05)])
def test_point_coeffs_match_oracle(ves, alpha):
    """get_pt_coeffs / get_pt_dcoeffs (src/uv_map.C:14-420): the weights
    reproduce the oracle's position and tangents, sum to one (position) and
    zero (derivatives), and list the face's control points."""
    w, g, o = ves
    r = w.r.copy()
    r[-3:] = alpha
    faces, uv = random_points(o, 3000, 11, deep=120)
    ok = well_conditioned(o, faces, uv)
    say("[points] coeffs n=%d (%d near a vertex) alpha=%s" % (len(faces), (~ok).sum(), alpha))
    ids, W = g.point_coeffs(faces, uv)
    ncp = (ids >= 0).sum(axis=1)
    val = np.array([len(x) for x in meshgen.read_mesh(w.mesh_path).rings])
    assert (ncp[faces < o.nreg] == 12).all()
    irr = faces >= o.nreg
    assert (ncp[irr] == val[ids[irr, 0]] + 6).all()
    assert (np.where(ids[:, None, :] < 0, W, 0) == 0).all()
    s = W.sum(axis=2)
    scale = np.abs(W).sum(axis=2)
    say("   sum w0 - 1", np.abs(s[:, 0] - 1).max(), "sum dw / sum |dw|", (np.abs(s[:, 1:]) / scale[:, 1:]).max())
    assert np.abs(s[:, 0] - 1).max() <= 1e-13
    # near the vertex the derivative weights are sums of 2^d-scaled entries
    # that cancel: the sum rule holds to the evaluation's conditioning there,
    # which the tangents below check against the oracle's own response
    assert (np.abs(s[ok, 1:]) <= 1e-13 * scale[ok, 1:]).all()
    Ro, _ = o.point_eval(r, faces, uv)
    R = _coupled(r, ids, W, 0)
    say("   R rel", rel(R, Ro))
    assert rel(R, Ro) <= 1e-12
    Go = o.point_geometry(r, faces, uv)[:, :6]
    T = np.concatenate([_coupled(r, ids, W, 1), _coupled(r, ids, W, 2)], 1)
    say("   Ru, Rv rel (well conditioned)", rel(T[ok], Go[ok]))
    assert rel(T[ok], Go[ok]) <= 1e-12
    sens = ulp_response(lambda q: o.point_geometry(r, faces[~ok], q)[:, :6], uv[~ok])
    err = np.abs(T[~ok] - Go[~ok])
    say("   near the vertex: max err", err.max(), "max 1-ulp response", sens.max())
    assert (err <= 1e-12 * np.abs(Go[ok]).max() + 1000 * sens).all()


def test_point_coeffs_at_extraordinary_vertex(ves):
    """The reference's mask at the vertex itself: 1/2 on the vertex, 1/(2 val)
    on its ring in the mesh's ring order, derivative weights zero."""
    w, g, o = ves
    rings = meshgen.read_mesh(w.mesh_path).rings
    faces = np.arange(o.nreg, o.nfaces, 7, dtype=np.int32)
    uv = np.zeros((faces.size, 2))
    ids, W = g.point_coeffs(faces, uv)
    for k in range(faces.size):
        i = ids[k, 0]
        val = len(rings[i])
        assert val != 6
        assert list(ids[k, 1:1 + val]) == list(rings[i])
        assert (ids[k, 1 + val:] == -1).all()
        assert W[k, 0, 0] == 0.5 and (W[k, 0, 1:1 + val] == 1.0 / (2 * val)).all()
        assert (W[k, 1:] == 0).all() and (W[k, 0, 1 + val:] == 0).all()
    R = _coupled(w.r, ids, W, 0)
    Ro, _ = o.point_eval(w.r, faces, uv)
    assert rel(R, Ro) <= 1e-12


@pytest.mark.parametrize("mesh", ["vesicle", "defect_sheet"])
def test_point_coeffs_couple_the_gradient(ves, mesh):
    """The gradient the reference's particle couplings build from the weights
    (dE/dr_x = alpha w_0x dE/dR; src/pcomplex.C:160-200) equals pointGradient's
    with no normal term, periodic images and valence-7 faces included."""
    if mesh == "vesicle":
        w, g, o = ves
    else:
        w = workloads.defect_sheet()
        g, o = _pair(w)
    r = w.r.copy()
    r[-3:] = (1.1, 0.9, 1.05)
    faces, uv = random_points(o, 1500, 12)
    say("[points] coeffs gradient coupling mesh=%s n=%d" % (mesh, faces.size))
    rng = np.random.default_rng(13)
    a = rng.normal(size=(faces.size, 3))
    G, _ = g.point_grad(r, faces, uv, a, np.zeros_like(a))
    ids, W = g.point_coeffs(faces, uv)
    nv = (r.size - 3) // 3
    Gc = np.zeros((nv, 3))
    m = ids >= 0
    np.add.at(Gc, ids[m], (W[:, 0][m])[:, None] * (a * r[-3:])[np.nonzero(m)[0]])
    say("   vertex gradient rel", rel(Gc.ravel(), G[:-3]))
    assert rel(Gc.ravel(), G[:-3]) <= 1e-12


def test_threaded_staging_matches_serial(ves):
    """Batches of 32768 points and more are resolved and staged over host
    threads: per-point results equal those of small (single-threaded)
    batches bitwise, and a rejection names the first bad point."""
    w, g, o = ves
    r = w.r.copy()
    r[-3:] = (1.1, 0.9, 1.05)
    n = 150000
    faces, uv = random_points(o, n, 21, deep=40)
    say("[points] threaded staging n=%d" % n)
    R, N = g.point_eval(r, faces, uv)
    Gm = g.point_geometry(r, faces, uv)
    ids, W = g.point_coeffs(faces, uv)
    rng = np.random.default_rng(22)
    a, b = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    _, P = g.point_grad(r, faces, uv, a, b)
    for k in range(0, n, 20000):
        s = slice(k, k + 20000)
        R2, N2 = g.point_eval(r, faces[s], uv[s])
        assert np.array_equal(R2, R[s]) and np.array_equal(N2, N[s], equal_nan=True)
        assert np.array_equal(g.point_geometry(r, faces[s], uv[s]), Gm[s], equal_nan=True)
        i2, W2 = g.point_coeffs(faces[s], uv[s])
        assert np.array_equal(i2, ids[s]) and np.array_equal(W2, W[s])
        _, P2 = g.point_grad(r, faces[s], uv[s], a[s], b[s])
        assert np.array_equal(P2, P[s])
    bad_f, bad_uv = faces.copy(), uv.copy()
    bad_f[120000] = o.nfaces
    bad_uv[70000] = np.nan
    with pytest.raises(S.UmcpError, match="point 70000: non-finite"):
        g.point_eval(r, bad_f, bad_uv)
    with pytest.raises(S.UmcpError, match="point 120000: face index out of range"):
        g.point_grad(r, bad_f, uv, a, b)
