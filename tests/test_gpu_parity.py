"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar (FP64 path): energy relative error <= 1e-12 and gradient norm-wise
relative error <= 1e-12 against the oracle on the same seeded inputs; the
per-triangle lipid packing is identical; Brownian trajectories driven by one
host-generated noise array agree to 1e-9 A after 25 steps (positions are
~1e2 A, so that is ~1e-11 relative).  Determinism: repeated evaluations are
bitwise equal.
"""
import sys

import numpy as np
import pytest

try:  # load torch's HIP runtime before libumcp_core.so so both share one
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from oracle import oracle as O
from umcp_amd import surface as S
from umcp_amd import workloads

pytestmark = pytest.mark.gpu

E_TOL = 1e-12
G_TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(w, c0=None, params=None):
    prm = dict(w.params)
    if params:
        prm.update(params)
    g = S.Surface(w.mesh_path)
    g.set_params(**prm)
    g.generatePlan()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**{k: v for k, v in prm.items() if k != "precision"})
    if c0 is not None:
        g.set_face_c0(c0)
        o.set_face_c0(c0)
    g.setg0(w.r0)
    o.setg0(w.r0)
    return g, o


def check_eval(g, o, r, tag, alpha_tol=1e-11, g_tol=G_TOL):
    say("[parity]", tag, "nv", g.nv, "faces", g.nfaces)
    E_o = o.energy(r)
    g_o = o.grad(r)
    E_g, g_g = g.energy_grad(r)
    rel_e = abs(E_g - E_o) / abs(E_o)
    rel_g = np.linalg.norm(g_g[:-3] - g_o[:-3]) / np.linalg.norm(g_o[:-3])
    rel_a = np.abs(g_g[-3:] - g_o[-3:]).max() / np.abs(g_o[-3:]).max()
    say("   E", E_g, E_o, "rel", rel_e, "grad rel", rel_g, "alpha rel", rel_a)
    assert rel_e <= E_TOL
    assert rel_g <= g_tol
    assert rel_a <= alpha_tol
    return E_g, g_g


@pytest.fixture(scope="module")
def ves642():
    return workloads.vesicle_642()


def test_vesicle_642_lipid_mode(ves642):
    g, o = make_pair(ves642)
    check_eval(g, o, ves642.r, "vesicle_642 mode lipid")
    for f in range(g.nfaces):
        t = o.face_info(f)["tri"]
        assert g.face_lipids(f) == o.tri_lipids(t)


# every FP64 launch shape: point split across 1-4 waves with or without the
# register bound, and the control-point split over 2-4 waves
SHAPES = [(s, f) for s in (1, 2, 3, 4) for f in (0, S.LAUNCH_OCC2)] + [(s, S.LAUNCH_CPS) for s in (2, 3, 4)]


def test_launch_shapes(ves642):
    """Every launch shape meets the bar."""
    g, o = make_pair(ves642)
    for split, flags in SHAPES:
        g.set_launch(split, flags)
        check_eval(g, o, ves642.r, "vesicle_642 split %d flags %d" % (split, flags))


GRAD_FORMS = [S.GRAD_REFERENCE, S.GRAD_EXACT]


@pytest.mark.parametrize("grad_form", GRAD_FORMS)
def test_all_terms_helfrich(grad_form):
    """Every term of the Helfrich form, in both gradient forms: surface::grad's
    (src/grad.C:448) and the exact derivative."""
    w = workloads.vesicle(4, 300.0, 1.5, seed=21, name="vesicle_2562_terms")
    g, o = make_pair(w, params=dict(kc=6.0, kg=-2.0, KA=0.5, k_reg=1.5, mode=1, grad_form=grad_form))
    c0 = workloads.face_c0_uniform(g.nfaces, 4, -0.02, 0.02)
    g.set_face_c0(c0)
    o.set_face_c0(c0)
    check_eval(g, o, w.r, "vesicle_2562 helfrich kg kreg KA grad_form %d" % grad_form)


def test_gradient_forms_differ():
    """The two Helfrich gradient forms give the same energy and gradients
    that differ by surface::grad's dE/dg term, by as much on the GPU as in the
    oracle."""
    w = workloads.vesicle(4, 300.0, 1.5, seed=21, name="vesicle_2562_terms")
    out = {}
    for form in GRAD_FORMS:
        g, o = make_pair(w, params=dict(kc=6.0, mode=1, grad_form=form))
        c0 = workloads.face_c0_uniform(g.nfaces, 4, -0.02, 0.02)
        g.set_face_c0(c0)
        o.set_face_c0(c0)
        out[form] = (g.energy_grad(w.r), o.grad(w.r))
    (E_r, g_r), o_r = out[S.GRAD_REFERENCE]
    (E_x, g_x), o_x = out[S.GRAD_EXACT]
    assert E_r == E_x
    d_g, d_o = g_r - g_x, o_r - o_x
    assert np.linalg.norm(d_o) > 1e-3 * np.linalg.norm(o_x)
    assert np.linalg.norm(d_g - d_o) <= 1e-11 * np.linalg.norm(d_o)


@pytest.mark.parametrize("form", ["lipid", "helfrich", "helfrich_exact", "sheet"])
def test_valence_5_and_7(form):
    """A sphere with valence-5 and valence-7 vertices (four edge flips on an
    icosphere, one Loop step): the valence-7 face class against the oracle,
    in every launch shape, with both restraints in the Helfrich form (both
    gradient forms); and a periodic sheet whose irregular faces reach across
    the box."""
    w = workloads.defect_sheet() if form == "sheet" else workloads.defect_vesicle()
    if form in ("lipid", "sheet"):
        g, o = make_pair(w)
    else:
        gf = S.GRAD_EXACT if form == "helfrich_exact" else S.GRAD_REFERENCE
        g, o = make_pair(w, params=dict(kc=6.0, kg=-2.0, KA=0.5, k_reg=1.5, mode=1, grad_form=gf))
        c0 = workloads.face_c0_uniform(g.nfaces, 6, -0.02, 0.02)
        g.set_face_c0(c0)
        o.set_face_c0(c0)
        rst = {"KA_global": 0.1, "KV": 0.004}
        g.set_restraints(**rst)
        o.set_restraints(**rst)
    vals = sorted(c["valence"] for c in g.classes())
    assert vals == [5, 6, 7], vals
    for split, flags in SHAPES:
        g.set_launch(split, flags)
        check_eval(g, o, w.r, "defect %s split %d flags %d" % (form, split, flags))


def test_periodic_sheet_small():
    w = workloads.sheet(32, Lx=250.0, seed=5, name="sheet_32_parity")
    g, o = make_pair(w)
    check_eval(g, o, w.r, "sheet_32 periodic")


def test_sheet_128(tmp_path):
    w = workloads.sheet_128()
    g, o = make_pair(w)
    check_eval(g, o, w.r, "sheet_128")


@pytest.mark.parametrize("grad_form", GRAD_FORMS)
def test_vesicle_40962_full_size(grad_form):
    w = workloads.vesicle_40962()
    g, o = make_pair(w, params=dict(mode=1, grad_form=grad_form))
    c0 = workloads.face_c0_uniform(g.nfaces, w.extra["c0_seed"])
    g.set_face_c0(c0)
    o.set_face_c0(c0)
    g.set_restraints(**w.extra["restraints"])
    o.set_restraints(**w.extra["restraints"])
    # With the area restraint, dE/dalpha is dominated by lamA * 2A, lamA =
    # 2 KA_g (A - A0)/A0.  A - A0 is ~4e-6 of A here, so the ~1e-13 relative
    # rounding of the two 1.2M-point area sums (GPU tree order vs the serial
    # oracle) becomes ~1e-8 in lamA: the alpha bound is that conditioning.
    check_eval(g, o, w.r, "vesicle_40962 helfrich c0, area and volume restraints, grad_form %d" % grad_form,
               alpha_tol=2e-8)


def test_grad_accumulates_and_is_deterministic(ves642):
    g, o = make_pair(ves642)
    r = ves642.r
    E1, g1 = g.energy_grad(r)
    E2, g2 = g.energy_grad(r)
    assert E1 == E2 and np.array_equal(g1, g2)
    acc = np.ones_like(r)
    g.grad(r, acc)
    np.testing.assert_array_equal(acc, 1.0 + g1)
    assert g.energy(r) == E1


def test_nan_energy_is_an_error(ves642):
    """A NaN position makes the energy NaN; the reference prints "nan en." and
    exits (src/interpLib.C:3546-3550).  Here the energy and gradient entry points return
    UMCP_ERR_NAN_ENERGY, leave the caller's gradient untouched, and the handle
    evaluates finite positions afterwards exactly as before."""
    g, o = make_pair(ves642)
    r = ves642.r
    E1, g1 = g.energy_grad(r)
    bad = r.copy()
    bad[3 * 17 + 1] = np.nan
    say("[parity] nan position at vertex 17")
    with pytest.raises(S.UmcpError, match=r"\(-2\).*nan energy"):
        g.energy(bad)
    acc = np.full_like(r, 2.5)
    with pytest.raises(S.UmcpError, match=r"\(-2\)"):
        g.grad(bad, acc)
    assert np.all(acc == 2.5)
    E2, g2 = g.energy_grad(r)
    assert E2 == E1 and np.array_equal(g2, g1)


def test_area(ves642):
    g, o = make_pair(ves642)
    A, A0 = g.area(ves642.r)
    Ao, A0o = o.area(ves642.r)
    say("[parity] area", A, Ao, A0, A0o)
    assert abs(A - Ao) <= 1e-12 * Ao and abs(A0 - A0o) <= 1e-12 * A0o


def test_bd_trajectory(ves642):
    g, o = make_pair(ves642)
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    xi = np.random.default_rng(99).standard_normal((25, 3 * g.nv))
    say("[parity] bd 25 steps", mob, amp)
    r_g, e_g = g.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    r_o, e_o = o.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    err = np.abs(r_g - r_o).max()
    say("   max |dr|", err, "energy rel", np.abs(e_g / e_o - 1).max())
    assert err <= 1e-9
    assert np.abs(e_g / e_o - 1).max() <= 1e-11


def test_bd_nan_energy_is_an_error(ves642):
    """Brownian steps that report energies stop with UMCP_ERR_NAN_ENERGY once a
    step's energy is NaN (the reference's src/interpLib.C:3546-3550 check, met
    through the energy it evaluates every step); the same handle then steps
    finite positions exactly as a fresh run does."""
    g, o = make_pair(ves642)
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    xi = np.random.default_rng(7).standard_normal((4, 3 * g.nv))
    r1, e1 = g.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    bad = ves642.r.copy()
    bad[3 * 300 + 2] = np.nan
    say("[parity] bd with a nan position at vertex 300")
    with pytest.raises(S.UmcpError, match=r"\(-2\).*nan energy"):
        g.bd_steps(bad, xi, mob, amp, want_energies=True)
    r2, e2 = g.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    assert np.array_equal(r2, r1) and np.array_equal(e2, e1)
    say("[parity] bd without energies, nan position")
    with pytest.raises(S.UmcpError, match=r"\(-2\).*nan energy"):
        g.bd_steps(bad, xi, mob, amp, want_energies=False)
    r3 = g.bd_steps(ves642.r, xi, mob, amp, want_energies=False)
    assert np.array_equal(r3, r1)


def test_device_paths_flag_nan_energy(ves642):
    """The device entry points return before the work runs; a non-finite
    energy on any of them is reported by nan_check (then cleared), for a
    surface and for the members of an ensemble."""
    if torch is None:
        pytest.skip("torch is needed for device buffers")
    g, o = make_pair(ves642)
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    bad = ves642.r.copy()
    bad[3 * 41] = np.nan
    d_bad = torch.tensor(bad, device="cuda", dtype=torch.float64)
    d_ok = torch.tensor(ves642.r, device="cuda", dtype=torch.float64)
    d_E = torch.zeros(1, device="cuda", dtype=torch.float64)
    d_g = torch.zeros_like(d_ok)
    xi = torch.zeros((2, 3 * g.nv), device="cuda", dtype=torch.float64)
    g.nan_check()
    say("[parity] device energy_grad with a nan position")
    g.energy_grad_dev(d_bad, d_E, d_g)
    with pytest.raises(S.UmcpError, match=r"\(-2\).*nan energy"):
        g.nan_check()
    g.nan_check()                        # cleared
    g.energy_grad_dev(d_ok, d_E, d_g)
    g.nan_check()                        # finite
    for path in ("bd_steps_dev", "bd_steps_rng_dev", "noalloc"):
        say("[parity] device path", path)
        d_r = d_bad.clone()
        if path == "bd_steps_dev":
            g.bd_steps_dev(d_r, xi, 2, mob, amp, None)
        elif path == "bd_steps_rng_dev":
            g.set_step(0)
            g.bd_steps_rng_dev(d_r, 2, mob, amp, 5, None)
        else:
            g.energy_grad_dev_noalloc(d_r, d_E, d_g)
        with pytest.raises(S.UmcpError, match=r"\(-2\)"):
            g.nan_check()
    say("[parity] ensemble member with a nan position")
    ens = S.Ensemble()
    h = make_pair(ves642)[0]
    h.set_step(0)
    g.set_step(0)
    d_a, d_b = d_ok.clone(), d_bad.clone()
    ens.add(g, d_a, mob, amp, 1, None)
    ens.add(h, d_b, mob, amp, 2, None)
    ens.bd_steps(2, 0)
    assert ens.nan_check() == 1
    assert ens.nan_check() is None
    ens.close()


def test_device_entry_points(ves642):
    if torch is None:
        pytest.skip("torch is needed for device buffers")
    g, o = make_pair(ves642)
    r = ves642.r
    E_h, g_h = g.energy_grad(r)
    d_r = torch.tensor(r, device="cuda", dtype=torch.float64)
    d_E = torch.zeros(1, device="cuda", dtype=torch.float64)
    d_g = torch.zeros_like(d_r)
    torch.cuda.synchronize()
    g.energy_grad_dev(d_r, d_E, d_g)   # enqueued on the surface's own stream
    torch.cuda.synchronize()           # device-wide: waits for that stream too
    assert float(d_E.cpu()[0]) == E_h
    np.testing.assert_array_equal(d_g.cpu().numpy(), g_h)


@pytest.mark.parametrize("case", ["vesicle_lipid", "defect_helfrich"])
def test_face_and_vertex_energies(case):
    """surface::faceEnergy of every face, energy(r, NULL, do_vertex) and
    area(r, do_vertex): each face's energy against the oracle's to 1e-12 of
    the largest; the faces' energies sum to the total energy and their area
    terms to surface::area's; a vertex's energy is the sum over the faces that
    have it as a control point (computeModifiedVertices' list, duplicates
    included), and its area the same faces' area terms; moving that vertex
    alone changes the total energy by exactly the change of its vertex energy
    (a face it moves but the list missed would break it)."""
    if case == "vesicle_lipid":
        w = workloads.vesicle_642()
        g, o = make_pair(w)
    else:
        w = workloads.defect_vesicle()
        g, o = make_pair(w, params=dict(kc=6.0, kg=-2.0, KA=0.5, k_reg=1.5, mode=1))
        c0 = workloads.face_c0_uniform(g.nfaces, 8, -0.02, 0.02)
        g.set_face_c0(c0)
        o.set_face_c0(c0)
    r = w.r
    Ef = g.face_energies(r)
    Eo = np.array([o.face_energy(f, r) for f in range(o.nfaces)])
    say("[parity] face energies", case, "max rel", np.abs(Ef - Eo).max() / np.abs(Eo).max())
    assert np.abs(Ef - Eo).max() <= 1e-12 * np.abs(Eo).max()
    E_tot = g.energy(r)
    assert abs(Ef.sum() - E_tot) <= 1e-12 * abs(E_tot)
    # the faces' area terms sum to surface::area's totals (oracle-checked)
    _, Af, A0f = g.face_energies(r, areas=True)
    At, A0t = o.area(r)
    assert abs(Af.sum() - At) <= 1e-12 * At and abs(A0f.sum() - A0t) <= 1e-12 * A0t
    faces_of = [[] for _ in range(o.nv)]
    for f in range(o.nfaces):
        for v in o.face_info(f)["cp"]:
            faces_of[v].append(f)
    rng = np.random.default_rng(41)
    for v in list(rng.choice(o.nv, 12, replace=False)) + [0, o.nv - 1]:
        Ev = g.energy_vertex(r, int(v))
        ref = sum(o.face_energy(f, r) for f in faces_of[v])
        assert abs(Ev - ref) <= 1e-12 * max(abs(ref), 1e-300) + 1e-14 * abs(E_tot), v
        r2 = r.copy()
        r2[3 * v:3 * v + 3] += rng.normal(0.0, 0.3, 3)
        dE = g.energy(r2) - E_tot
        dEv = g.energy_vertex(r2, int(v)) - Ev
        assert abs(dE - dEv) <= 1e-11 * abs(E_tot), (v, dE, dEv)
        # area(r, do_vertex): the same faces' area and reference area
        A, A0 = g.area_vertex(r, int(v))
        assert abs(A - Af[faces_of[v]].sum()) <= 1e-13 * A and abs(A0 - A0f[faces_of[v]].sum()) <= 1e-13 * A0
    with pytest.raises(S.UmcpError):
        g.energy_vertex(r, o.nv)


def _morton_group_span(mesh, group=64):
    """The most distinct control-point vertices (corners and their rings) of
    any run of `group` triangles taken in Morton order of their centroids,
    the order the plan groups faces in."""
    tris = mesh.triangles()
    c = mesh.r[tris].mean(axis=1)
    lo, hi = c.min(0), c.max(0)
    q = ((c - lo) / (hi - lo + 1e-12) * 1023).astype(np.int64)

    def spread(x):
        x = x & 0x3FF
        x = (x | (x << 16)) & 0x030000FF
        x = (x | (x << 8)) & 0x0300F00F
        x = (x | (x << 4)) & 0x030C30C3
        return (x | (x << 2)) & 0x09249249

    order = np.argsort(spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2), kind="stable")
    best = 0
    for s in range(0, len(order), group):
        vs = set()
        for t in tris[order[s:s + group]]:
            for v in t:
                vs.add(int(v))
                vs.update(mesh.rings[int(v)])
        best = max(best, len(vs))
    return best


def test_scattered_geometry(tmp_path):
    """An icosphere's topology with its vertex positions dealt out at random:
    faces that are close in space are far apart on the mesh, so a 64-face
    group spans more distinct vertices than the kernels stage in LDS (192)
    and the face kernel and the restraint pre-pass read their control points
    directly.  Every launch shape against the oracle (positive energy terms
    only, kg = 0, and the global area restraint).  The energy bar is the
    usual 1e-12; the gradient's is 1e-11: the scattered faces fold through
    one another, the energy is ~1e3 times a vesicle's, and the gradient is a
    sum of large terms of both signs (measured 1.2e-12).  A wrong index on
    this path gives errors of order one."""
    from umcp_amd import meshgen
    mesh = meshgen.icosphere(3, 120.0)
    mesh.r = mesh.r[np.random.default_rng(17).permutation(mesh.nv)]
    assert _morton_group_span(mesh) > 192
    path = str(tmp_path / "scattered_geometry.mesh")
    meshgen.write_mesh(mesh, path)
    r0 = meshgen.flat_coords(mesh)
    r = r0.copy()
    r[:-3] += np.random.default_rng(23).normal(0.0, 0.3, 3 * mesh.nv)
    w = workloads.Workload(path, path, mesh.nv, r0, r, dict(kc=6.0, kg=0.0, KA=0.5, k_reg=1.5, mode=1), T=298.0)
    g, o = make_pair(w)
    c0 = workloads.face_c0_uniform(g.nfaces, 8, -0.02, 0.02)
    g.set_face_c0(c0)
    o.set_face_c0(c0)
    g.set_restraints(KA_global=0.1)
    o.set_restraints(KA_global=0.1)
    for split, flags in SHAPES:
        g.set_launch(split, flags)
        check_eval(g, o, w.r, "scattered geometry split %d flags %d" % (split, flags), g_tol=1e-11)


@pytest.mark.parametrize("case", ["vesicle", "sheet"])
def test_box_scale_not_one(case):
    """The box scale alpha (the last three entries of r) away from 1: the
    face kernel scales every contracted derivative by it, divides it out of
    the scattered gradient and forms dE/dalpha from the unscaled positions.
    The vesicle with per-face c0 and both restraints (whose volume seed is
    scaled the same way), the periodic defect sheet, every launch shape,
    against the oracle at the usual bars."""
    if case == "vesicle":
        w = workloads.vesicle_642()
        g, o = make_pair(w, params=dict(kc=6.0, kg=-2.0, KA=0.5, k_reg=1.5, mode=1))
        c0 = workloads.face_c0_uniform(g.nfaces, 12, -0.02, 0.02)
        g.set_face_c0(c0)
        o.set_face_c0(c0)
        rst = {"KA_global": 0.1, "KV": 0.004}
        g.set_restraints(**rst)
        o.set_restraints(**rst)
    else:
        w = workloads.defect_sheet()
        g, o = make_pair(w)
    r = w.r.copy()
    r[-3:] = [1.02, 0.97, 1.01]
    for split, flags in SHAPES:
        g.set_launch(split, flags)
        check_eval(g, o, r, "alpha %s %s split %d flags %d" % (list(r[-3:]), case, split, flags))
