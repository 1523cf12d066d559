"""GPU: the device-resident point evaluation and gradient
(umcp_point_eval_dev, umcp_point_grad_dev) against the host-array calls
(umcp_point_eval, umcp_point_grad), which the points tests pin against the
oracle (evaluateRNRM, src/interpLib.C:7936-8196; pointGradient, src/grad.C:
5697-6443).  Both resolve the points by
the same rules and evaluate them with the same kernels, so the results are
bitwise equal, also for sub-patch levels 1-40 next to an extraordinary
vertex (random points do not fall within rounding of a level border, where
the device's log10 could pick the adjacent level), the vertex itself, a
periodic sheet and a million points on the 40962-vertex vesicle; the
gradient with frac_mult, the on_surface term and accumulation into the
caller's arrays.  Rejected points get NaN outputs (evaluation) or add
nothing (gradient), leave the others unchanged and are named by
umcp_point_status."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(w):
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    return g


def random_points(nfaces, nreg, n, seed, deep=0, vertex=0):
    rng = np.random.default_rng(seed)
    faces = rng.integers(0, nfaces, n)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    uv = np.stack([u, v], 1)
    irr = np.flatnonzero(faces >= nreg)
    for k, i in enumerate(irr[:deep]):
        s = 10.0 ** -(1 + k % 12)
        uv[i] *= s / max(uv[i].sum(), 1e-300)
    for i in irr[deep:deep + vertex]:
        uv[i] = 0.0
    return faces.astype(np.int32), uv


def dev_eval(g, r, faces, uv):
    d_r = torch.tensor(r, device="cuda", dtype=torch.float64)
    d_f = torch.tensor(faces, device="cuda", dtype=torch.int32)
    d_uv = torch.tensor(uv, device="cuda", dtype=torch.float64)
    d_R = torch.zeros((len(faces), 3), device="cuda", dtype=torch.float64)
    d_N = torch.zeros_like(d_R)
    g.point_eval_dev(d_r, d_f, d_uv, d_R, d_N)
    torch.cuda.synchronize()
    return d_R.cpu().numpy(), d_N.cpu().numpy()


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("name", ["vesicle_642", "defect_sheet"])
def test_device_equals_host(name):
    w = {"vesicle_642": workloads.vesicle_642, "defect_sheet": lambda: workloads.defect_sheet()}[name]()
    g = make(w)
    o = O.OracleSurface(w.mesh_path)
    r = w.r.copy()
    r[-3:] = [1.02, 0.97, 1.01]     # alpha != 1
    faces, uv = random_points(o.nfaces, o.nreg, 20000, 3, deep=600, vertex=4)
    say("[points dev]", name, "points", len(faces), "irregular", int((faces >= o.nreg).sum()))
    R_h, N_h = g.point_eval(r, faces, uv)
    R_d, N_d = dev_eval(g, r, faces, uv)
    g.point_status()
    assert same(R_d, R_h) and same(N_d, N_h)
    # repeatable
    R_2, N_2 = dev_eval(g, r, faces, uv)
    assert same(R_2, R_d) and same(N_2, N_d)


def test_rejections_name_the_first_point():
    w = workloads.vesicle_642()
    g = make(w)
    o = O.OracleSurface(w.mesh_path)
    faces, uv = random_points(o.nfaces, o.nreg, 5000, 8)
    R_h, N_h = g.point_eval(w.r, faces, uv)
    irr = int(np.flatnonzero(faces >= o.nreg)[0])
    bad_face = faces.copy()
    bad_face[4000] = o.nfaces + 3              # face out of range
    bad_uv = uv.copy()
    bad_uv[irr] = [0.9, 0.9]                    # outside an irregular face
    bad_uv[4500, 0] = np.nan                    # non-finite
    R_d, N_d = dev_eval(g, w.r, bad_face, bad_uv)
    with pytest.raises(S.UmcpError) as ex:
        g.point_status()
    first = min(4000, irr, 4500)
    say("[points dev] rejection:", ex.value)
    assert ("point %d:" % first) in str(ex.value)
    bad = np.zeros(len(faces), bool)
    bad[[4000, irr, 4500]] = True
    assert np.isnan(R_d[bad]).all() and np.isnan(N_d[bad]).all()
    assert same(R_d[~bad], R_h[~bad]) and same(N_d[~bad], N_h[~bad])
    # the host call rejects the same batch, naming the same point
    with pytest.raises(S.UmcpError) as ex2:
        g.point_eval(w.r, bad_face, bad_uv)
    assert ("point %d:" % first) in str(ex2.value)
    # a clean batch clears the status
    dev_eval(g, w.r, faces, uv)
    g.point_status()


def test_million_points_full_size():
    w = workloads.vesicle_40962()
    g = make(w)
    faces, uv = random_points(g.nfaces, g.nreg, 1_000_000, 11, deep=60)
    R_h, N_h = g.point_eval(w.r, faces, uv)
    R_d, N_d = dev_eval(g, w.r, faces, uv)
    g.point_status()
    say("[points dev] 1M points equal:", same(R_d, R_h), same(N_d, N_h))
    assert same(R_d, R_h) and same(N_d, N_h)


def dev_grad(g, r, faces, uv, a, b, frac=None, on_surface=False, kT=0.0, g0=None, gp0=None):
    t = lambda x, dt=torch.float64: torch.tensor(x, device="cuda", dtype=dt)  # noqa: E731
    d_g = t(g0 if g0 is not None else np.zeros(len(r)))
    d_gp = t(gp0 if gp0 is not None else np.zeros((len(faces), 2)))
    g.point_grad_dev(t(r), t(faces, torch.int32), t(uv), t(a), t(b), d_g, d_gp,
                     d_frac=t(frac) if frac is not None else None, on_surface=on_surface, kT=kT)
    torch.cuda.synchronize()
    return d_g.cpu().numpy(), d_gp.cpu().numpy()


@pytest.mark.parametrize("name", ["vesicle_642", "defect_sheet"])
def test_gradient_device_equals_host(name):
    w = {"vesicle_642": workloads.vesicle_642, "defect_sheet": lambda: workloads.defect_sheet()}[name]()
    g = make(w)
    o = O.OracleSurface(w.mesh_path)
    r = w.r.copy()
    r[-3:] = [1.02, 0.97, 1.01]
    faces, uv = random_points(o.nfaces, o.nreg, 20000, 5, deep=600)
    rng = np.random.default_rng(6)
    a, b = rng.normal(size=(len(faces), 3)), rng.normal(size=(len(faces), 3))
    frac = rng.uniform(0.2, 1.0, len(faces))
    g_start = rng.normal(size=len(r))
    gp_start = rng.normal(size=(len(faces), 2))
    g_h, gp_h = g.point_grad(r, faces, uv, a, b, frac_mult=frac, on_surface=True, kT=0.6, g=g_start.copy(),
                             g_puv=gp_start.copy())
    g_d, gp_d = dev_grad(g, r, faces, uv, a, b, frac, True, 0.6, g_start, gp_start)
    g.point_status()
    say("[points dev] gradient", name, "equal:", np.array_equal(g_d, g_h), np.array_equal(gp_d, gp_h))
    assert np.array_equal(g_d, g_h) and np.array_equal(gp_d, gp_h)
    # without frac and on_surface, twice in a row (the work space reused)
    g_h2, gp_h2 = g.point_grad(r, faces[:777], uv[:777], a[:777], b[:777])
    g_d2, gp_d2 = dev_grad(g, r, faces[:777], uv[:777], a[:777], b[:777])
    assert np.array_equal(g_d2, g_h2) and np.array_equal(gp_d2, gp_h2)


def test_gradient_rejection_adds_nothing_for_that_point():
    w = workloads.vesicle_642()
    g = make(w)
    o = O.OracleSurface(w.mesh_path)
    faces, uv = random_points(o.nfaces, o.nreg, 3000, 9)
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=(len(faces), 3)), rng.normal(size=(len(faces), 3))
    irr = int(np.flatnonzero(faces >= o.nreg)[0])
    bad_uv = uv.copy()
    bad_uv[irr] = 0.0                      # the extraordinary vertex: no gradient rule
    keep = np.ones(len(faces), bool)
    keep[irr] = False
    g_h, gp_h = g.point_grad(w.r, faces[keep], uv[keep], a[keep], b[keep])
    g_d, gp_d = dev_grad(g, w.r, faces, bad_uv, a, b)
    with pytest.raises(S.UmcpError) as ex:
        g.point_status()
    assert ("point %d:" % irr) in str(ex.value)
    assert np.array_equal(g_d, g_h) and not gp_d[irr].any()
    assert np.array_equal(gp_d[keep], gp_h)


def test_gradient_million_points_full_size():
    w = workloads.vesicle_40962()
    g = make(w)
    faces, uv = random_points(g.nfaces, g.nreg, 1_000_000, 13, deep=60)
    rng = np.random.default_rng(14)
    a, b = rng.normal(size=(len(faces), 3)), rng.normal(size=(len(faces), 3))
    g_h, gp_h = g.point_grad(w.r, faces, uv, a, b, on_surface=True, kT=0.6)
    g_d, gp_d = dev_grad(g, w.r, faces, uv, a, b, on_surface=True, kT=0.6)
    g.point_status()
    say("[points dev] 1M gradient equal:", np.array_equal(g_d, g_h), np.array_equal(gp_d, gp_h))
    assert np.array_equal(g_d, g_h) and np.array_equal(gp_d, gp_h)


@pytest.mark.parametrize("name", ["vesicle_642", "defect_sheet"])
def test_geometry_device_equals_host(name):
    w = {"vesicle_642": workloads.vesicle_642, "defect_sheet": lambda: workloads.defect_sheet()}[name]()
    g = make(w)
    o = O.OracleSurface(w.mesh_path)
    r = w.r.copy()
    r[-3:] = [1.02, 0.97, 1.01]
    faces, uv = random_points(o.nfaces, o.nreg, 20000, 17, deep=600)
    G_h = g.point_geometry(r, faces, uv)
    d_G = torch.zeros((len(faces), 15), device="cuda", dtype=torch.float64)
    g.point_geometry_dev(torch.tensor(r, device="cuda"), torch.tensor(faces, device="cuda", dtype=torch.int32),
                         torch.tensor(uv, device="cuda"), d_G)
    torch.cuda.synchronize()
    g.point_status()
    assert same(d_G.cpu().numpy(), G_h)
    # the extraordinary vertex has no geometry: NaN record, named
    irr = int(np.flatnonzero(faces >= o.nreg)[0])
    uv2 = uv.copy()
    uv2[irr] = 0.0
    g.point_geometry_dev(torch.tensor(r, device="cuda"), torch.tensor(faces, device="cuda", dtype=torch.int32),
                         torch.tensor(uv2, device="cuda"), d_G)
    torch.cuda.synchronize()
    with pytest.raises(S.UmcpError) as ex:
        g.point_status()
    assert ("point %d:" % irr) in str(ex.value)
    G2 = d_G.cpu().numpy()
    assert np.isnan(G2[irr]).all()
    keep = np.ones(len(faces), bool)
    keep[irr] = False
    assert same(G2[keep], G_h[keep])
