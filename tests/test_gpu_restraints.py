"""GPU parity of the global restraints and of the device noise path.

Bars: FP64 energy and gradient with area and volume restraints <= 1e-12
relative to the oracle (targets away from the current state so both
multipliers are non-zero); volume and area values <= 1e-12; Brownian
trajectories with restraints, one host noise array, <= 1e-9 A after 10 steps;
the device noise equals the CPU restatement of Philox4x64-10 + Box-Muller to
1e-13 (libm-level differences only); device-noise steps equal explicit-noise
steps fed the same deviates bitwise; FP32-mixed with restraints on the
40962-vertex vesicle within 1e-6 (energy) and 1e-5 (gradient RMS); runs of
several Brownian steps bitwise equal to one-step calls, and on the
40962-vertex vesicle (restraint finalize as a launch) within 1e-9 A of the
oracle.
"""
import sys

import numpy as np
import pytest

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

from oracle import oracle as O
from umcp_amd import surface as S
from umcp_amd import workloads

import noise_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(w, params=None, c0=None, restraints=None):
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
    if restraints:
        g.set_restraints(**restraints)
        o.set_restraints(**restraints)
    return g, o


@pytest.fixture(scope="module")
def ves642():
    return workloads.vesicle_642()


def away_targets(o, r, KA_global=0.4, KV=0.02):
    A, _ = o.area(r)
    V = o.volume(r)
    return dict(KA_global=KA_global, KV=KV, A0=0.97 * A, V0=1.04 * V)


def test_volume_and_area_values(ves642):
    g, o = make_pair(ves642)
    say("[restraints] volume/area 642")
    V_g, V_o = g.volume(ves642.r), o.volume(ves642.r)
    A_g, A0_g = g.area(ves642.r)
    A_o, A0_o = o.area(ves642.r)
    say("   V", V_g, V_o, "A", A_g, A_o)
    assert abs(V_g - V_o) <= 1e-12 * abs(V_o)
    assert abs(A_g - A_o) <= 1e-12 * A_o


@pytest.mark.parametrize("mode", [0, 1])
def test_restraint_energy_gradient(ves642, mode):
    o0 = O.OracleSurface(ves642.mesh_path)
    rs = away_targets(o0, ves642.r)
    c0 = np.linspace(-0.02, 0.02, 1280)
    g, o = make_pair(ves642, params=dict(mode=mode), c0=c0, restraints=rs)
    say("[restraints] 642 mode", mode, rs)
    E_o, g_o = o.energy(ves642.r), o.grad(ves642.r)
    E_g, g_g = g.energy_grad(ves642.r)
    e_r, la, lv = o.restraint_energy(ves642.r)
    rel_e = abs(E_g - E_o) / abs(E_o)
    rel_g = np.linalg.norm(g_g[:-3] - g_o[:-3]) / np.linalg.norm(g_o[:-3])
    say("   E", E_g, E_o, "restraint part", e_r, "lamA", la, "lamV", lv, "rel", rel_e, rel_g)
    assert abs(la) > 1e-3 and abs(lv) > 1e-5
    assert rel_e <= 1e-12
    assert rel_g <= 1e-12
    assert np.abs(g_g[-3:] - g_o[-3:]).max() <= 1e-11 * np.abs(g_o[-3:]).max()


def test_restraints_vanish_at_setg0(ves642):
    g, o = make_pair(ves642)
    E_plain = g.energy(ves642.r0)
    g.set_restraints(KA_global=0.4, KV=0.02)
    E_rs = g.energy(ves642.r0)
    say("[restraints] setg0 state", E_plain, E_rs)
    assert abs(E_rs - E_plain) <= 1e-12 * max(1.0, abs(E_plain))


def test_periodic_sheet_rejects_volume(mesh_dir):
    w = workloads.sheet(32, seed=4, name="sheet32_rs")
    g = S.Surface(w.mesh_path)
    g.set_params(**w.params)
    g.generatePlan()
    g.setg0(w.r0)
    with pytest.raises(S.UmcpError):
        g.set_restraints(KV=0.01)
    # the area restraint is defined on periodic meshes
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    A, _ = o.area(w.r)
    g.set_restraints(KA_global=0.3, A0=0.99 * A)
    o.set_restraints(KA_global=0.3, A0=0.99 * A)
    E_g, g_g = g.energy_grad(w.r)
    E_o, g_o = o.energy(w.r), o.grad(w.r)
    say("[restraints] sheet area", E_g, E_o)
    assert abs(E_g - E_o) <= 1e-12 * abs(E_o)
    assert np.linalg.norm(g_g[:-3] - g_o[:-3]) <= 1e-12 * np.linalg.norm(g_o[:-3])


def test_bd_with_restraints(ves642):
    o0 = O.OracleSurface(ves642.mesh_path)
    g, o = make_pair(ves642, restraints=away_targets(o0, ves642.r, 0.4, 0.05))
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    xi = np.random.default_rng(5).standard_normal((10, 3 * g.nv))
    say("[restraints] bd 10 steps")
    r_g, e_g = g.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    r_o, e_o = o.bd_steps(ves642.r, xi, mob, amp, want_energies=True)
    err = np.abs(r_g - r_o).max()
    say("   max |dr|", err)
    assert err <= 1e-9
    assert np.abs(e_g / e_o - 1).max() <= 1e-11


@pytest.mark.skipif(torch is None, reason="torch provides the device buffers")
def test_device_noise_matches_restatement():
    n = 700
    d = torch.zeros(3 * n, device="cuda", dtype=torch.float64)
    for seed, step in ((12345, 0), (2 ** 62 + 11, 4999)):
        say("[noise] seed", seed, "step", step)
        S.noise(seed, step, n, d)
        torch.cuda.synchronize()
        ref = noise_ref.normals(seed, step, n)
        err = np.abs(d.cpu().numpy() - ref).max()
        say("   max err", err)
        assert err <= 1e-13


@pytest.mark.skipif(torch is None, reason="torch provides the device buffers")
def test_rng_steps_equal_explicit_steps(ves642):
    o0 = O.OracleSurface(ves642.mesh_path)
    rs = away_targets(o0, ves642.r, 0.4, 0.05)
    a, _ = make_pair(ves642, restraints=rs)
    b, _ = make_pair(ves642, restraints=rs)
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    K, seed = 6, 777
    say("[noise] rng vs explicit", K)
    d_ra = torch.tensor(ves642.r, device="cuda", dtype=torch.float64)
    d_rb = d_ra.clone()
    e_a = torch.zeros(K, device="cuda", dtype=torch.float64)
    e_b = torch.zeros(K, device="cuda", dtype=torch.float64)
    xi = torch.zeros((K, 3 * a.nv), device="cuda", dtype=torch.float64)
    a.set_step(0)
    a.bd_steps_rng_dev(d_ra, K, mob, amp, seed, e_a)
    for k in range(K):
        S.noise(seed, k, a.nv, xi[k])
    torch.cuda.synchronize()
    b.bd_steps_dev(d_rb, xi, K, mob, amp, e_b)
    torch.cuda.synchronize()
    assert a.get_step() == K
    assert torch.equal(d_ra, d_rb)
    assert torch.equal(e_a, e_b)


def test_mixed_with_restraints_40962():
    w = workloads.vesicle_40962()
    c0 = workloads.face_c0_uniform(81920, w.extra["c0_seed"])
    o0 = O.OracleSurface(w.mesh_path)
    o0.set_params(**w.params)
    rs = away_targets(o0, w.r, 0.2, 0.01)
    g, o = make_pair(w, params=dict(precision=S.PREC_MIXED), c0=c0, restraints=rs)
    say("[restraints] mixed 40962")
    E_o, g_o = o.energy(w.r), o.grad(w.r)
    E_g, g_g = g.energy_grad(w.r)
    rel_e = abs(E_g - E_o) / abs(E_o)
    rms = lambda x: float(np.sqrt(np.mean(x ** 2)))  # noqa: E731
    rel_g = rms(g_g[:-3] - g_o[:-3]) / rms(g_o[:-3])
    say("   energy rel", rel_e, "grad rms rel", rel_g)
    assert rel_e <= 1e-6
    assert rel_g <= 1e-5


@pytest.mark.skipif(torch is None, reason="torch provides the device buffers")
@pytest.mark.parametrize("mesh", ["vesicle_642", "vesicle_40962"])
def test_step_runs_equal_single_steps(mesh):
    """K Brownian steps in one call and K calls of one step give the same
    bits (642 vertices: restraint finalize folded into the face blocks;
    40962: a finalize launch)."""
    w = workloads.vesicle_642() if mesh == "vesicle_642" else workloads.vesicle_40962()
    o0 = O.OracleSurface(w.mesh_path)
    o0.set_params(**w.params)
    rs = away_targets(o0, w.r, 0.3, 0.02)
    c0 = workloads.face_c0_uniform(81920, w.extra["c0_seed"]) if mesh == "vesicle_40962" else None
    a, _ = make_pair(w, c0=c0, restraints=rs)
    b, _ = make_pair(w, c0=c0, restraints=rs)
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    K = 5
    say("[steps] %s: %d steps in one call against one-step calls" % (mesh, K))
    xi = torch.tensor(np.random.default_rng(9).standard_normal((K, 3 * a.nv)), device="cuda", dtype=torch.float64)
    d_ra = torch.tensor(w.r, device="cuda", dtype=torch.float64)
    d_rb = d_ra.clone()
    e_a = torch.zeros(K, device="cuda", dtype=torch.float64)
    e_b = torch.zeros(K, device="cuda", dtype=torch.float64)
    a.bd_steps_dev(d_ra, xi, K, mob, amp, e_a)
    for k in range(K):
        b.bd_steps_dev(d_rb, xi[k:k + 1], 1, mob, amp, e_b[k:k + 1])
    torch.cuda.synchronize()
    say("   max |dr|", float((d_ra - d_rb).abs().max()))
    assert torch.equal(d_ra, d_rb)
    assert torch.equal(e_a, e_b)


def test_bd_with_restraints_40962():
    """Three Brownian steps of the 40962-vertex vesicle (per-face c0, both
    restraints, targets away from the start) against the oracle."""
    w = workloads.vesicle_40962()
    c0 = workloads.face_c0_uniform(81920, w.extra["c0_seed"])
    o0 = O.OracleSurface(w.mesh_path)
    o0.set_params(**w.params)
    g, o = make_pair(w, c0=c0, restraints=away_targets(o0, w.r, 0.3, 0.02))
    mob, amp = S.bd_coefficients(1e-12, 1e-9, 298.0)
    xi = np.random.default_rng(6).standard_normal((3, 3 * g.nv))
    say("[restraints] bd 40962 3 steps")
    r_g, e_g = g.bd_steps(w.r, xi, mob, amp, want_energies=True)
    r_o, e_o = o.bd_steps(w.r, xi, mob, amp, want_energies=True)
    err = np.abs(r_g - r_o).max()
    say("   max |dr|", err, "energies", e_g, e_o)
    assert err <= 1e-9
    assert np.abs(e_g / e_o - 1).max() <= 1e-11
