"""The oracle's effective-mass integrators (src/hd.C, src/modes.C), pinned.

- mass matrix: exactly symmetric; partition of unity of the limit-surface
  position coefficients at every lattice point makes the sum of all entries
  ntri * 66 * m; positive definite;
- the irregular evaluation next to the extraordinary vertex tends to the
  vertex's limit mask (the special case the reference uses at u = v = 0);
- Cholesky inverse: M Mi = I and agreement with numpy.linalg.inv;
- integrators: with Mi = c I the Brownian form reduces to the per-vertex
  Brownian step (mobility c Delta^2, noise c sqrt(2 kT Delta / gamma)); with
  zero force and zero temperature the Langevin momenta decay by exactly
  (1 - gamma Delta c) per step; at vanishing friction and zero temperature the
  Langevin form conserves E + 1/2 p.Mi.p (the check of the reference's own
  energy_conservation test, here without particles).
"""
import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import units, workloads

AKMA = 1.0 / (0.04888e-12)


@pytest.fixture(scope="module")
def ves():
    w = workloads.vesicle_642()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.setg0(w.r0)
    return w, o


def test_mass_matrix_partition_of_unity(ves):
    w, o = ves
    m = o.effm_point_mass()
    M = o.effm_mass()
    assert np.array_equal(M, M.T)
    assert abs(M.sum() - o.ntri() * 66 * m) <= 1e-12 * M.sum()
    assert np.linalg.eigvalsh(M).min() > 0


@pytest.mark.parametrize("val", [5, 7])
def test_irregular_evaluation_tends_to_limit_mask(val):
    tab = np.zeros(6 * (val + 6))
    O.lib().orc_irregular_table(val, 1e-9, 1e-9, O._p(tab))
    mask = np.zeros(val + 6)
    mask[0] = 0.5
    mask[1:1 + val] = 1.0 / (2 * val)
    assert np.abs(tab[:val + 6] - mask).max() < 1e-6
    assert abs(tab[:val + 6].sum() - 1.0) < 1e-12


def test_inverse_and_truncation(ves):
    w, o = ves
    M = o.effm_mass()
    Mi = o.effm_invert(M)
    assert np.abs(M @ Mi - np.eye(o.nv)).max() < 1e-10
    ref = np.linalg.inv(M)
    assert np.abs(Mi - ref).max() <= 1e-12 * np.abs(ref).max()
    d = np.sqrt(np.abs(np.outer(np.diag(Mi), np.diag(Mi))))
    for tol in (1e-20, 1e-4):
        T, kept = o.effm_truncate(Mi, tol)
        keep = np.abs(Mi) / d > tol
        assert kept == int(keep.sum())
        assert np.array_equal(T, np.where(keep, Mi, 0.0))


def test_brownian_form_reduces_to_per_vertex_step(ves):
    w, o = ves
    c, dt, kT, gamma = 4.7e-7, 1e-11, units.kT(298.0), 0.001
    Mi = c * np.eye(o.nv)
    xi = np.random.default_rng(5).standard_normal((4, 3 * o.nv))
    r_e, _ = o.effm_steps(w.r, np.zeros(3 * o.nv), xi, 0, dt, kT, gamma, Mi)
    D = AKMA * dt
    r_b = o.bd_steps(w.r, xi, c * D * D, c * np.sqrt(2 * gamma * kT * D) / gamma)
    step = np.abs(r_b - w.r).max()
    assert np.abs(r_e - r_b).max() <= 1e-10 * step


def test_langevin_friction_decay():
    w = workloads.vesicle_642()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(kc=0.0, KA=0.0, kg=0.0, k_reg=0.0, mode=O.MODE_LIPID)
    o.setg0(w.r0)
    c, dt, gamma = 5e-7, 1e-11, 0.3
    D = AKMA * dt
    pp0 = np.random.default_rng(2).standard_normal(3 * o.nv)
    xi = np.zeros((6, 3 * o.nv))
    r, pp = o.effm_steps(w.r, pp0, xi, 1, dt, 0.0, gamma, c * np.eye(o.nv))
    f = 1.0 - gamma * D * c
    assert np.allclose(pp, pp0 * f ** 6, rtol=1e-12, atol=0)
    drift = c * D * sum(f ** k for k in range(1, 7))
    # positions near 100 A carry about 1e-14 A of rounding each step
    assert np.abs((r[:-3] - w.r[:-3]) - drift * pp0).max() <= 1e-12


def test_langevin_conserves_energy_without_friction(ves):
    w, o = ves
    Mi = o.effm_inverse()
    dt, n = 2e-13, 200
    xi = np.zeros((n, 3 * o.nv))
    r, pp, e = o.effm_steps(w.r, np.zeros(3 * o.nv), xi, 1, dt, 0.0, 1e-12, Mi, energies=True)
    kin = 0.5 * sum(pp[c::3] @ Mi @ pp[c::3] for c in range(3))
    H0, H1 = e[0], o.energy(r) + kin
    moved = np.abs(r - w.r).max()
    assert moved > 0.05            # the membrane has moved
    assert abs(H1 - H0) <= 2e-3 * abs(H0)
