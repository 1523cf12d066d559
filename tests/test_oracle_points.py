"""The oracle's points on the limit surface (surface::evaluateRNRM,
src/interpLib.C:7936-8196; surface::pointGradient, src/grad.C:5697-6443),
pinned by what the mathematics fixes rather than by retyping its formulas:

- corners: R at the corners of every face is the Loop limit point of the
  corner vertex (regular vertices: mask 1/2, 1/12; src/interpLib.C:997-1098
  uses the same points), and the owner sits at (u, v) = (0, 0);
- Gauss points: R at the 15 regular quadrature points equals the plan's
  tables contracted with the positions (the tables are pinned against the
  reference's quadrature separately);
- derivatives: g_puv with de_dr = e_k is (Ru_k, Rv_k), and with de_dnrm = e_k
  it is (Nu_k, Nv_k); both match central differences of the evaluated R and
  normal in u and v, on regular faces and on every sub-patch (three
  sub-domains, several levels) of irregular faces;
- vertex gradient: the gradient of de_dr.R + de_dnrm.nrm over every control
  point coordinate matches central differences, and so does the alpha
  gradient of de_dr.R (regular faces, alpha != 1); on_surface adds the
  gradient of -kT log g; frac_mult scales the vertex gradient only;
- the reference's rejections: the extraordinary vertex itself for the
  gradient (a NULL map there), levels beyond MAX_DOMAIN = 100;
- point geometry (surface::ru, rv, g, c): the tangents are central
  differences of R; with alpha != 1 everything but g equals the geometry of
  alpha-scaled positions at alpha = 1; g is |Ru x Rv| on regular faces and
  the reference's u v / 4^level scaling of it on irregular ones; the
  Weingarten map from central differences of the normal gives c1 + c2, c1 c2
  and the principal directions (its eigenvectors), which map to orthogonal
  tangents.
"""
import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import workloads

REG_PTS = [(0.1, 0.1), (0.3, 0.1), (0.5, 0.1), (0.7, 0.1), (0.9, 0.1), (0.1, 0.3), (0.3, 0.3),
           (0.5, 0.3), (0.7, 0.3), (0.1, 0.5), (0.3, 0.5), (0.5, 0.5), (0.1, 0.7), (0.3, 0.7),
           (0.1, 0.9)]
# irregular parameters covering the three sub-domains at levels 1, 3 and deep
IRR_UV = [(0.7, 0.2), (0.2, 0.7), (0.3, 0.3), (0.15, 0.05), (0.01, 0.13), (0.07, 0.08),
          (3e-6, 1e-6), (0.5, 0.5)]
REG_UV = [(0.2, 0.3), (0.05, 0.9), (0.6, 0.1), (0.0, 0.0), (1.0, 0.0)]


@pytest.fixture(scope="module")
def ves():
    w = workloads.vesicle_642()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    return w, o


def _faces(o):
    reg = [0, 7, o.nreg // 2, o.nreg - 1]
    irr = [o.nreg, o.nreg + 17, o.nreg + 33, o.nfaces - 1]
    return reg, irr


def test_corners_are_limit_points(ves):
    w, o = ves
    r = w.r.copy()
    L = o.limit_points(r).reshape(-1, 3)
    scale = np.abs(L).max()
    for f in range(o.nfaces):
        info = o.face_info(f)
        corners = [(1.0, 0.0), (0.0, 1.0)] + ([(0.0, 0.0)] if f < o.nreg else [])
        R, _ = o.point_eval(r, [f] * len(corners), corners)
        for k, x in enumerate(R):
            d = np.abs(L[info["cp"]] - x).max(axis=1)
            assert d.min() < 1e-12 * scale, (f, corners[k])
        if f < o.nreg:
            assert np.abs(R[2] - L[info["owner"]]).max() < 1e-12 * scale


def test_regular_gauss_points_match_tables(ves):
    w, o = ves
    r = w.r.copy()
    r[-3:] = [1.1, 0.9, 1.05]
    nv = o.nv
    for f in _faces(o)[0]:
        info = o.face_info(f)
        R, _ = o.point_eval(r, [f] * 15, REG_PTS)
        for p in range(15):
            w6, _, pbc, _ = o.face_table(f, p)
            q = r[:3 * nv].reshape(-1, 3)[info["cp"]] + pbc
            ref = (w6[0][:, None] * q).sum(axis=0) * r[-3:]
            assert np.abs(R[p] - ref).max() < 1e-12 * np.abs(ref).max()


def _fd_uv(o, r, f, u, v, h=1e-6):
    Rp, Np = o.point_eval(r, [f, f, f, f], [(u + h, v), (u - h, v), (u, v + h), (u, v - h)])
    return ((Rp[0] - Rp[1]) / (2 * h), (Rp[2] - Rp[3]) / (2 * h),
            (Np[0] - Np[1]) / (2 * h), (Np[2] - Np[3]) / (2 * h))


@pytest.mark.parametrize("kind", ["regular", "irregular"])
def test_uv_derivatives_match_differences(ves, kind):
    w, o = ves
    r = w.r.copy()
    reg, irr = _faces(o)
    faces, uvs = (reg, REG_UV[:3]) if kind == "regular" else (irr, IRR_UV[:6])
    for f in faces:
        for (u, v) in uvs:
            Ru_fd, Rv_fd, Nu_fd, Nv_fd = _fd_uv(o, r, f, u, v)
            for k in range(3):
                e = np.zeros((1, 3))
                e[0, k] = 1.0
                _, gp = o.point_grad(r, [f], [(u, v)], e, np.zeros((1, 3)))
                s = max(np.abs(Ru_fd).max(), np.abs(Rv_fd).max())
                assert abs(gp[0, 0] - Ru_fd[k]) < 1e-6 * s and abs(gp[0, 1] - Rv_fd[k]) < 1e-6 * s
                _, gp = o.point_grad(r, [f], [(u, v)], np.zeros((1, 3)), e)
                s = max(np.abs(Nu_fd).max(), np.abs(Nv_fd).max(), 1e-3)
                assert abs(gp[0, 0] - Nu_fd[k]) < 1e-5 * s and abs(gp[0, 1] - Nv_fd[k]) < 1e-5 * s


def _point_energy(o, r, f, uv, a, b):
    R, N = o.point_eval(r, [f], [uv])
    return float(a @ R[0] + b @ N[0])


@pytest.mark.parametrize("kind", ["regular", "irregular"])
def test_vertex_gradient_matches_differences(ves, kind):
    w, o = ves
    r = w.r.copy()
    rng = np.random.default_rng(11)
    reg, irr = _faces(o)
    cases = [(reg[1], (0.2, 0.3)), (reg[2], (0.6, 0.1))] if kind == "regular" else \
        [(irr[0], (0.3, 0.3)), (irr[1], (0.01, 0.13)), (irr[2], (0.7, 0.2))]
    for f, uv in cases:
        a, b = rng.normal(size=3), rng.normal(size=3)
        g, _ = o.point_grad(r, [f], [uv], a[None], b[None])
        cp = o.face_info(f)["cp"]
        h = 1e-4
        for vtx in cp:
            for k in range(3):
                rp, rm = r.copy(), r.copy()
                rp[3 * vtx + k] += h
                rm[3 * vtx + k] -= h
                fd = (_point_energy(o, rp, f, uv, a, b) - _point_energy(o, rm, f, uv, a, b)) / (2 * h)
                assert abs(g[3 * vtx + k] - fd) < 1e-6 * max(1.0, abs(fd)), (f, vtx, k)
        # nothing outside the control points
        mask = np.ones(o.nv, bool)
        mask[cp] = False
        assert not g[:3 * o.nv].reshape(-1, 3)[mask].any()


def test_alpha_gradient_matches_differences(ves):
    w, o = ves
    r = w.r.copy()
    r[-3:] = [1.1, 0.9, 1.05]
    reg, _ = _faces(o)
    a = np.array([0.3, -1.2, 0.7])
    for f in reg[:2]:
        g, _ = o.point_grad(r, [f], [(0.25, 0.35)], a[None], np.zeros((1, 3)))
        for k in range(3):
            h = 1e-7
            rp, rm = r.copy(), r.copy()
            rp[3 * o.nv + k] += h
            rm[3 * o.nv + k] -= h
            fd = (_point_energy(o, rp, f, (0.25, 0.35), a, np.zeros(3)) -
                  _point_energy(o, rm, f, (0.25, 0.35), a, np.zeros(3))) / (2 * h)
            assert abs(g[3 * o.nv + k] - fd) < 1e-6 * abs(fd)


def _log_g(o, r, f, uv):
    """log of the area element |Ru x Rv| from the evaluated tangents."""
    e = np.eye(3)
    Ru, Rv = np.zeros(3), np.zeros(3)
    for k in range(3):
        _, gp = o.point_grad(r, [f], [uv], e[k][None], np.zeros((1, 3)))
        Ru[k], Rv[k] = gp[0]
    return np.log(np.linalg.norm(np.cross(Ru, Rv)))


@pytest.mark.parametrize("kind", ["regular", "irregular"])
def test_on_surface_term_is_entropy_gradient(ves, kind):
    w, o = ves
    r = w.r.copy()
    reg, irr = _faces(o)
    f, uv = (reg[0], (0.3, 0.2)) if kind == "regular" else (irr[3], (0.15, 0.05))
    kT = 0.616
    z = np.zeros((1, 3))
    g0, _ = o.point_grad(r, [f], [uv], z, z)
    g1, gp = o.point_grad(r, [f], [uv], z, z, on_surface=1, kT=kT)
    assert not g0.any() and not gp.any()
    cp = o.face_info(f)["cp"]
    h = 1e-4
    for vtx in cp[:6]:
        for k in range(3):
            rp, rm = r.copy(), r.copy()
            rp[3 * vtx + k] += h
            rm[3 * vtx + k] -= h
            fd = -kT * (_log_g(o, rp, f, uv) - _log_g(o, rm, f, uv)) / (2 * h)
            assert abs(g1[3 * vtx + k] - fd) < 1e-6 * max(1e-3, abs(fd))


def test_frac_mult_scales_vertex_gradient_only(ves):
    w, o = ves
    r = w.r.copy()
    reg, irr = _faces(o)
    faces = [reg[1], irr[1]]
    uv = [(0.2, 0.3), (0.3, 0.3)]
    a = np.array([[0.1, 0.2, 0.3], [-0.4, 0.5, 0.6]])
    b = np.array([[0.7, -0.8, 0.9], [1.0, 1.1, -1.2]])
    g1, p1 = o.point_grad(r, faces, uv, a, b)
    g2, p2 = o.point_grad(r, faces, uv, a, b, frac_mult=[0.5, 0.5])
    assert np.allclose(g2, 0.5 * g1, rtol=1e-14, atol=0)
    assert np.array_equal(p1, p2)


def test_reference_rejections(ves):
    w, o = ves
    r = w.r.copy()
    _, irr = _faces(o)
    z = np.zeros((1, 3))
    with pytest.raises(ValueError):
        o.point_grad(r, [irr[0]], [(0.0, 0.0)], z, z)
    with pytest.raises(ValueError):
        o.point_eval(r, [irr[0]], [(1e-40, 1e-40)])
    # the vertex itself evaluates to the reference's mask with a NaN normal
    R, N = o.point_eval(r, [irr[0]], [(0.0, 0.0)])
    i = o.face_info(irr[0])["owner"]
    nb = o.neighbors(i)
    q = r[:3 * o.nv].reshape(-1, 3)
    ref = 0.5 * q[i] + q[nb].sum(axis=0) / (2 * len(nb))
    assert np.abs(R[0] - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.isnan(N).all()


# ---------------------------------------------------------------------------
# point geometry (surface::ru, rv, g, c; src/uv_map.C:884-2990)
# ---------------------------------------------------------------------------

GEOM_CASES = [("regular", 1, (0.2, 0.3)), ("regular", 2, (0.6, 0.1)), ("irregular", 0, (0.3, 0.3)),
              ("irregular", 1, (0.01, 0.13)), ("irregular", 2, (0.7, 0.2)), ("irregular", 3, (0.15, 0.05))]


def _case(o, kind, k):
    reg, irr = _faces(o)
    return (reg if kind == "regular" else irr)[k]


def test_geometry_tangents_are_derivatives(ves):
    """Ru, Rv are central differences of the evaluated (alpha-scaled) R."""
    w, o = ves
    r = w.r.copy()
    r[-3:] = [1.1, 0.9, 1.05]
    for kind, k, (u, v) in GEOM_CASES:
        f = _case(o, kind, k)
        G = o.point_geometry(r, [f], [(u, v)])[0]
        Ru_fd, Rv_fd, _, _ = _fd_uv(o, r, f, u, v)
        s = np.abs(Ru_fd).max() + np.abs(Rv_fd).max()
        assert np.abs(G[0:3] - Ru_fd).max() < 1e-6 * s and np.abs(G[3:6] - Rv_fd).max() < 1e-6 * s, (kind, k)


def test_geometry_alpha_is_a_scaling_of_positions(ves):
    """Everything but g is the geometry of the surface through alpha-scaled
    positions with alpha = 1 (g, like the reference's, ignores alpha)."""
    w, o = ves
    al = np.array([1.1, 0.9, 1.05])
    r = w.r.copy()
    r[-3:] = al
    r1 = np.concatenate([(w.r[:-3].reshape(-1, 3) * al).reshape(-1), [1.0, 1.0, 1.0]])
    for kind, k, uv in GEOM_CASES:
        f = _case(o, kind, k)
        A = o.point_geometry(r, [f], [uv])[0]
        B = o.point_geometry(r1, [f], [uv])[0]
        assert np.allclose(A[[0, 1, 2, 3, 4, 5, 7, 8, 9, 10]], B[[0, 1, 2, 3, 4, 5, 7, 8, 9, 10]], rtol=1e-12,
                           atol=1e-14), (kind, k)
        G1 = o.point_geometry(w.r, [f], [uv])[0]
        assert A[6] == G1[6]


def test_geometry_area_element(ves):
    """g = |Ru x Rv| on regular faces; on irregular faces the reference's
    scaling by the face's own u v over the squared sub-patch factor 4^level."""
    w, o = ves
    r = w.r.copy()
    for kind, k, (u, v) in GEOM_CASES:
        f = _case(o, kind, k)
        G = o.point_geometry(r, [f], [(u, v)])[0]
        gn = np.linalg.norm(np.cross(G[0:3], G[3:6]))
        if kind == "regular":
            assert abs(G[6] - gn) < 1e-12 * gn
        else:
            level = int(np.ceil(-np.log2(u + v))) if u + v < 1 - 1e-9 else 1
            assert abs(G[6] - gn * u * v / 4.0 ** level) < 1e-12 * gn * u * v / 4.0 ** level, (k, level)


def test_geometry_curvature_from_the_normal(ves):
    """With alpha = 1 the Weingarten map from central differences of the
    evaluated normal, W = -(J^T J)^-1 J^T [n_u n_v] with J = [Ru Rv], gives
    c1 + c2 = -tr W and c1 c2 = det W; vec1 and vec2 are its eigenvectors for
    -c1 and -c2 and map to orthogonal tangent directions."""
    w, o = ves
    r = w.r.copy()
    for kind, k, (u, v) in GEOM_CASES:
        f = _case(o, kind, k)
        G = o.point_geometry(r, [f], [(u, v)])[0]
        _, _, Nu, Nv = _fd_uv(o, r, f, u, v, h=1e-5)
        J = np.stack([G[0:3], G[3:6]], 1)
        W = -np.linalg.solve(J.T @ J, J.T @ np.stack([Nu, Nv], 1))
        c1, c2 = G[9], G[10]
        sc = abs(c1) + abs(c2)
        assert abs(G[7] - (c1 + c2)) < 1e-14 * sc and abs(G[8] - c1 * c2) < 1e-14 * sc * sc
        assert abs(-np.trace(W) - G[7]) < 1e-6 * sc, (kind, k)
        assert abs(np.linalg.det(W) - G[8]) < 1e-6 * sc * sc, (kind, k)
        v1, v2 = G[11:13], G[13:15]
        assert abs(np.linalg.norm(v1) - 1) < 1e-14 and abs(np.linalg.norm(v2) - 1) < 1e-14
        assert np.abs(W @ v1 + c1 * v1).max() < 1e-5 * sc, (kind, k)
        assert np.abs(W @ v2 + c2 * v2).max() < 1e-5 * sc, (kind, k)
        t1, t2 = J @ v1, J @ v2
        assert abs(t1 @ t2) < 1e-6 * np.linalg.norm(t1) * np.linalg.norm(t2)


def test_geometry_rejections(ves):
    w, o = ves
    _, irr = _faces(o)
    with pytest.raises(ValueError):
        o.point_geometry(w.r, [irr[0]], [(0.0, 0.0)])
    with pytest.raises(ValueError):
        o.point_geometry(w.r, [irr[0]], [(1e-40, 1e-40)])
