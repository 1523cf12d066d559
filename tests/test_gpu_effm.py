"""GPU: the momentum-space integrators with the effective-mass matrix
(src/hd.C:1543-1720; src/modes.C:1719-1960) against the oracle.

Bars: the point mass and the truncated inverse effective mass equal the
oracle's to 1e-12 relative (two Cholesky implementations, same matrix);
Brownian and Langevin trajectories driven by the same host noise array match
the oracle's to 1e-9 of the displacement over ten steps (FP64 throughout; the
only differences are summation orders); the device noise stream gives the
trajectory the restated Philox normals give as a host array.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import units, workloads  # noqa: E402

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import noise_ref  # noqa: E402

pytestmark = pytest.mark.gpu

DT, GAMMA = 1e-11, 0.001


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pair():
    w = workloads.vesicle_642()
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    nnz, m = g.effm_build()
    return w, g, o, nnz, m, o.effm_inverse()


def test_inverse_mass_matches_oracle(pair):
    w, g, o, nnz, m, Mi_o = pair
    say("[effm] nnz", nnz, "point mass", m)
    assert abs(m - o.effm_point_mass()) <= 1e-12 * m
    Mi = g.effm_get()
    assert nnz == int((Mi_o != 0).sum())
    err = np.abs(Mi - Mi_o).max() / np.abs(Mi_o).max()
    say("   max rel diff", err)
    assert err <= 1e-12


@pytest.mark.parametrize("mode", [S.EFFM_BD, S.EFFM_LD])
def test_trajectory_matches_oracle(pair, mode):
    w, g, o, nnz, m, Mi_o = pair
    kT = units.kT(298.0)
    xi = np.random.default_rng(17 + mode).standard_normal((10, 3 * g.nv))
    pp0 = np.zeros(3 * g.nv)
    r_g, pp_g, e_g = g.effm_steps(w.r, pp0, xi, mode, DT, kT, GAMMA, energies=True)
    r_o, pp_o, e_o = o.effm_steps(w.r, pp0, xi, mode, DT, kT, GAMMA, Mi_o, energies=True)
    disp = np.abs(r_o - w.r).max()
    err = np.abs(r_g - r_o).max()
    say("[effm] mode", mode, "max displacement", disp, "max |dr|", err, "pp err", np.abs(pp_g - pp_o).max())
    assert err <= 1e-9 * disp
    assert np.abs(pp_g - pp_o).max() <= 1e-9 * max(np.abs(pp_o).max(), 1e-300) or mode == S.EFFM_BD
    if mode == S.EFFM_BD:
        assert not pp_g.any()           # momenta cleared after every Brownian step
    assert np.allclose(e_g, e_o, rtol=1e-11, atol=0)


def test_device_noise_equals_host_array(pair):
    w, g, o, nnz, m, Mi_o = pair
    kT = units.kT(298.0)
    steps, seed = 6, 99
    g.set_step(0)
    d_r = torch.tensor(w.r, device="cuda", dtype=torch.float64)
    d_pp = torch.zeros(3 * g.nv, device="cuda", dtype=torch.float64)
    g.effm_steps_dev(d_r, d_pp, steps, S.EFFM_LD, DT, kT, GAMMA, seed=seed)
    torch.cuda.synchronize()
    xi = np.stack([noise_ref.normals(seed, k, g.nv) for k in range(steps)])
    r_h, pp_h = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_LD, DT, kT, GAMMA)
    dr = np.abs(d_r.cpu().numpy() - r_h).max()
    say("[effm] device noise vs host array max |dr|", dr)
    assert dr <= 1e-9 * np.abs(r_h - w.r).max()


def test_larger_mesh_against_numpy_inverse():
    """2562 vertices (beyond the 642 the reference's dense CPU inverse is
    timed at): the device inverse against numpy's inverse of the oracle's mass
    matrix, and a Langevin trajectory against the oracle driven by that
    inverse."""
    w = workloads.vesicle(4, 200.0, 0.5, seed=31, name="effm_ves2562")
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    nnz, m = g.effm_build()
    Mi_ref, kept = o.effm_truncate(np.linalg.inv(o.effm_mass()))
    Mi = g.effm_get()
    err = np.abs(Mi - Mi_ref).max() / np.abs(Mi_ref).max()
    say("[effm] 2562: nnz", nnz, "of", g.nv ** 2, "max rel diff vs numpy", err)
    assert err <= 1e-11
    assert np.array_equal(Mi, Mi.T)
    kT = units.kT(298.0)
    xi = np.random.default_rng(23).standard_normal((4, 3 * g.nv))
    r_g, pp_g = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_LD, DT, kT, GAMMA)
    r_o, pp_o = o.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_LD, DT, kT, GAMMA, Mi_ref)
    disp = np.abs(r_o - w.r).max()
    say("   trajectory max |dr|", np.abs(r_g - r_o).max(), "displacement", disp)
    assert np.abs(r_g - r_o).max() <= 1e-9 * disp


def test_valence_7_mesh():
    """The sphere with valence-5 and valence-7 vertices: the mass matrix's
    lattice points on valence-7 faces (and the vertex mask at their corner)
    give the oracle's inverse, and a Brownian trajectory matches."""
    w = workloads.defect_vesicle()
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    nnz, m = g.effm_build()
    assert abs(m - o.effm_point_mass()) <= 1e-12 * m
    Mi_o = o.effm_inverse()
    Mi = g.effm_get()
    err = np.abs(Mi - Mi_o).max() / np.abs(Mi_o).max()
    say("[effm] valence 5/7 sphere: nnz", nnz, "max rel diff", err)
    assert err <= 1e-12 and nnz == int((Mi_o != 0).sum())
    kT = units.kT(298.0)
    xi = np.random.default_rng(29).standard_normal((6, 3 * g.nv))
    r_g, pp_g = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_BD, DT, kT, GAMMA)
    r_o, pp_o = o.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_BD, DT, kT, GAMMA, Mi_o)
    disp = np.abs(r_o - w.r).max()
    assert np.abs(r_g - r_o).max() <= 1e-9 * disp


def test_full_size_vesicle_builds_and_steps():
    """The 40962-vertex vesicle: a 13.4 GB inverse in HBM (far beyond the
    reference's CPU set-up), symmetric with a positive diagonal, and Langevin
    steps that stay finite."""
    w = workloads.vesicle_40962()
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    nnz, m = g.effm_build()
    say("[effm] 40962: nnz", nnz, "point mass", m)
    assert 0 < nnz <= g.nv ** 2
    # spot check of the kept (CSR) rows: symmetric pairs, positive diagonal
    Mi_rows = g.effm_get()
    for i in (0, 7, 20000, 40961):
        assert Mi_rows[i, i] > 0
        j = int(np.argmax(np.abs(Mi_rows[i]) * (np.arange(g.nv) != i)))
        assert Mi_rows[i, j] == Mi_rows[j, i]
    del Mi_rows
    kT = units.kT(298.0)
    d_r = torch.tensor(w.r, device="cuda", dtype=torch.float64)
    d_pp = torch.zeros(3 * g.nv, device="cuda", dtype=torch.float64)
    e = torch.zeros(3, device="cuda", dtype=torch.float64)
    g.set_step(0)
    g.effm_steps_dev(d_r, d_pp, 3, S.EFFM_LD, 1e-12, kT, GAMMA, seed=3, d_energies=e)
    torch.cuda.synchronize()
    assert torch.isfinite(d_r).all() and torch.isfinite(e).all()
    moved = float((d_r - torch.tensor(w.r, device="cuda", dtype=torch.float64)).abs().max())
    say("   3 steps, max displacement", moved)
    assert 0 < moved < 10.0


def test_periodic_sheet_against_oracle():
    """A periodic 32 x 32 sheet (1024 vertices, faces across the box
    boundary): inverse and a Brownian trajectory against the oracle."""
    w = workloads.sheet(32, seed=6, name="effm_sheet32")
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    nnz, m = g.effm_build()
    assert abs(m - o.effm_point_mass()) <= 1e-12 * m
    Mi_o = o.effm_inverse()
    Mi = g.effm_get()
    err = np.abs(Mi - Mi_o).max() / np.abs(Mi_o).max()
    say("[effm] sheet 32: nnz", nnz, "max rel diff", err)
    assert err <= 1e-11
    kT = units.kT(w.T)
    xi = np.random.default_rng(41).standard_normal((5, 3 * g.nv))
    r_g, _ = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_BD, DT, kT, GAMMA)
    r_o, _ = o.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_BD, DT, kT, GAMMA, Mi_o)
    disp = np.abs(r_o - w.r).max()
    say("   trajectory max |dr|", np.abs(r_g - r_o).max(), "displacement", disp)
    assert np.abs(r_g - r_o).max() <= 1e-9 * disp
