"""GPU: the scalable effective mass (umcp_effm_build_sparse): M kept sparse
and Mi v by a fixed-length Chebyshev solve, against the oracle's truncated
dense inverse (src/modes.C:1719-1960).

Bars: products Mi V equal the oracle's truncated inverse times V to 1e-11
(norm-wise; the truncation drops entries below 1e-20 of the diagonal, the
solve stops below a 1e-12 residual); at 10242 vertices, where the dense
inverse is too slow for the oracle, the residual |M x - v| / |v| formed with
the oracle's mass matrix is below 1e-12; trajectories of both integrators
match the oracle (or the dense form) to 1e-9 of the displacement; beyond the
dense form's vertex limit (163842 vertices) the solve is a linear, symmetric
operator (x.Mi y = y.Mi x) and the Langevin steps run.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import units, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

DT, GAMMA = 1e-11, 0.001


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(w, oracle=True):
    g = S.Surface(w.mesh_path)
    g.set_params(precision=S.PREC_FP64, **w.params)
    g.generatePlan()
    g.setg0(w.r0)
    o = None
    if oracle:
        o = O.OracleSurface(w.mesh_path)
        o.set_params(**w.params)
        o.setg0(w.r0)
    return g, o


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", ["vesicle_642", "defect_vesicle", "sheet32"])
def test_products_and_trajectories_match_oracle(name):
    w = {"vesicle_642": workloads.vesicle_642,
         "defect_vesicle": lambda: workloads.defect_vesicle(),
         "sheet32": lambda: workloads.sheet(32, seed=6, name="effm_sheet32")}[name]()
    g, o = make(w)
    nnz, m, iters = g.effm_build_sparse()
    M = o.effm_mass()
    say("[effm sparse]", name, "nv", g.nv, "nnz", nnz, "iterations", iters)
    assert nnz == int((M != 0).sum())
    assert abs(m - o.effm_point_mass()) <= 1e-12 * m
    Mi_o = o.effm_inverse()
    V = np.random.default_rng(5).standard_normal((g.nv, 3))
    X = g.effm_apply(V)
    err = rel(X, Mi_o @ V)
    say("   product rel err", err, "residual", np.linalg.norm(M @ X - V) / np.linalg.norm(V))
    assert err <= 1e-11
    kT = units.kT(w.T)
    for mode in (S.EFFM_BD, S.EFFM_LD):
        xi = np.random.default_rng(17 + mode).standard_normal((6, 3 * g.nv))
        pp0 = np.zeros(3 * g.nv)
        r_g, pp_g, e_g = g.effm_steps(w.r, pp0, xi, mode, DT, kT, GAMMA, energies=True)
        r_o, pp_o, e_o = o.effm_steps(w.r, pp0, xi, mode, DT, kT, GAMMA, Mi_o, energies=True)
        disp = np.abs(r_o - w.r).max()
        say("   mode", mode, "max |dr|", np.abs(r_g - r_o).max(), "displacement", disp)
        assert np.abs(r_g - r_o).max() <= 1e-9 * disp
        if mode == S.EFFM_LD:
            assert np.abs(pp_g - pp_o).max() <= 1e-9 * np.abs(pp_o).max()
        assert np.allclose(e_g, e_o, rtol=1e-11, atol=0)


def test_sparse_equals_dense_form_on_device():
    w = workloads.vesicle(4, 200.0, 0.5, seed=31, name="effm_ves2562")
    g, _ = make(w, oracle=False)
    kT = units.kT(298.0)
    xi = np.random.default_rng(23).standard_normal((4, 3 * g.nv))
    g.effm_build()
    r_d, pp_d = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_LD, DT, kT, GAMMA)
    V = np.random.default_rng(3).standard_normal((g.nv, 3))
    X_d = g.effm_apply(V)
    g.effm_build_sparse()
    with pytest.raises(S.UmcpError):
        g.effm_get()             # the sparse form keeps no inverse
    r_s, pp_s = g.effm_steps(w.r, np.zeros(3 * g.nv), xi, S.EFFM_LD, DT, kT, GAMMA)
    X_s = g.effm_apply(V)
    disp = np.abs(r_d - w.r).max()
    say("[effm sparse] 2562 dense vs sparse product", rel(X_s, X_d), "trajectory", np.abs(r_s - r_d).max(), disp)
    assert rel(X_s, X_d) <= 1e-11
    assert np.abs(r_s - r_d).max() <= 1e-9 * disp


def test_residual_at_10242_vertices():
    w = workloads.vesicle(5, 400.0, 0.5, seed=8, name="effm_ves10242")
    g, o = make(w)
    tol = 1e-12
    nnz, m, iters = g.effm_build_sparse(tol)
    M = o.effm_mass()
    V = np.random.default_rng(9).standard_normal((g.nv, 3))
    X = g.effm_apply(V)
    res = np.linalg.norm(M @ X - V) / np.linalg.norm(V)
    say("[effm sparse] 10242: iterations", iters, "residual", res)
    assert res <= tol


def test_beyond_dense_limit_163842():
    w = workloads.vesicle(7, 1600.0, 0.5, seed=12, name="effm_ves163842")
    g, _ = make(w, oracle=False)
    with pytest.raises(S.UmcpError):
        g.effm_build()            # the dense inverse is capped
    nnz, m, iters = g.effm_build_sparse()
    say("[effm sparse] 163842: nnz", nnz, "iterations", iters)
    rng = np.random.default_rng(4)
    U, V = rng.standard_normal((g.nv, 3)), rng.standard_normal((g.nv, 3))
    MU, MV = g.effm_apply(U), g.effm_apply(V)
    MW = g.effm_apply(2.0 * U - 0.5 * V)
    lin = rel(MW, 2.0 * MU - 0.5 * MV)
    sym = abs(np.sum(U * MV) - np.sum(V * MU)) / np.sqrt(np.sum(U * MU) * np.sum(V * MV))
    say("   linearity", lin, "symmetry", sym)
    assert lin <= 1e-13 and sym <= 1e-12
    assert np.sum(U * MU) > 0 and np.sum(V * MV) > 0
    kT = units.kT(298.0)
    d_r = torch.tensor(w.r, device="cuda", dtype=torch.float64)
    d_pp = torch.zeros(3 * g.nv, device="cuda", dtype=torch.float64)
    g.set_step(0)
    g.effm_steps_dev(d_r, d_pp, 3, S.EFFM_LD, DT, kT, GAMMA, seed=7)
    torch.cuda.synchronize()
    moved = (d_r[:-3].cpu().numpy() - w.r[:-3])
    assert np.isfinite(moved).all() and 0 < np.abs(moved).max() < 10.0


def test_sparse_device_noise_equals_host_array():
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    import noise_ref
    w = workloads.vesicle_642()
    g, _ = make(w, oracle=False)
    g.effm_build_sparse()
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
    say("[effm sparse] device noise vs host array max |dr|", dr)
    assert dr <= 1e-9 * np.abs(r_h - w.r).max()
    assert np.array_equal(d_pp.cpu().numpy(), pp_h) or np.abs(d_pp.cpu().numpy() - pp_h).max() <= 1e-9 * np.abs(pp_h).max()
