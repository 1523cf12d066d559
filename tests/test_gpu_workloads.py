"""The batched ensemble and the serving stream at their specified sizes, each
member or request tied to the oracle directly.

Batched ensemble (all 256 vesicles of `workloads.ensemble()`, levels 2-5,
Helfrich form, each with its own time step): 20 Brownian steps with device
noise, captured in one hipGraph and replayed, in FP64 and in FP32-mixed.
Bars:
* every level-5 member and 18 others (six each of levels 2, 3 and 4) end
  bitwise where they end alone (umcp_bd_steps_rng_dev), energies included;
* one member of each level, the level-5 one among them, against the oracle's
  Brownian steps (orc_bd_steps) fed the same deviates from the restated noise
  stream (tests/noise_ref.py): FP64 positions within 1e-9 A and energies
  within 1e-10 relative at every step; mixed energies within 1e-6 relative
  and positions within 1e-5 of the largest displacement (the FP32 budget of
  the gradient).

Serving stream: a mix holding a level-6 icosphere (40962 vertices) and a
256 x 256 periodic sheet (65536 vertices), each for 500 or more steps, beside
small requests, served with fewer slots than requests and chunks shorter
than every request.  Bars: every request bitwise equal to the request alone;
two small requests (levels 2 and 3, 500+ steps) against the oracle fed the
restated noise, FP64 positions within 1e-9 A.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import noise_ref  # noqa: E402
from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu

STEPS = 20
LAUNCH = (4, S.LAUNCH_OCC2)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _member_surface(m, precision):
    s = S.Surface(m.mesh_path)
    s.set_params(precision=precision, **m.params)
    s.generatePlan()
    s.set_launch(*LAUNCH)
    s.setg0(m.r0)
    s.set_step(0)
    return s


def _alone(m, precision, steps):
    s = _member_surface(m, precision)
    d_r = torch.tensor(m.r, device="cuda", dtype=torch.float64)
    e = torch.zeros(steps, device="cuda", dtype=torch.float64)
    mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
    s.bd_steps_rng_dev(d_r, steps, mob, amp, m.seed, e)
    torch.cuda.synchronize()
    return d_r.cpu().numpy(), e.cpu().numpy()


def _oracle_steps(mesh_path, params, r0, r, seed, steps, mob, amp, c0=None):
    o = O.OracleSurface(mesh_path)
    o.set_params(**params)
    if c0 is not None:
        o.set_face_c0(np.full(o.nfaces, c0))
    o.setg0(r0)
    xi = np.concatenate([noise_ref.normals(seed, k, o.nv) for k in range(steps)])
    return o.bd_steps(r, xi, mob, amp, want_energies=True)


@pytest.fixture(scope="module")
def members(mesh_dir):
    ms = workloads.ensemble()
    assert len(ms) == 256 and sorted({m.level for m in ms}) == [2, 3, 4, 5]
    return ms


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_MIXED])
def test_ensemble_256_members(members, precision):
    say("[c4] building 256 members, precision", precision)
    surfs = [_member_surface(m, precision) for m in members]
    d_rs = [torch.tensor(m.r, device="cuda", dtype=torch.float64) for m in members]
    d_es = [torch.zeros(STEPS, device="cuda", dtype=torch.float64) for _ in members]
    ens = S.Ensemble()
    ens.set_launch(LAUNCH[0], LAUNCH[1], 0, 0)
    for m, s, d_r, d_e in zip(members, surfs, d_rs, d_es):
        mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
        ens.add(s, d_r, mob, amp, m.seed, d_e)
    torch.cuda.synchronize()
    say("[c4] 20 steps in one captured graph")
    ens.bd_steps(STEPS, STEPS)
    torch.cuda.synchronize()
    got = [(d_r.cpu().numpy(), d_e.cpu().numpy()) for d_r, d_e in zip(d_rs, d_es)]
    ens.close()
    del surfs, d_rs, d_es

    pick = [i for i, m in enumerate(members) if m.level == 5]
    for L in (2, 3, 4):
        pick += [i for i, m in enumerate(members) if m.level == L][:6]
    say("[c4] alone runs:", len(pick), "members,", sum(members[i].level == 5 for i in pick), "of level 5")
    for i in pick:
        ra, ea = _alone(members[i], precision, STEPS)
        r, e = got[i]
        assert np.array_equal(r, ra), (i, members[i].level, float(np.abs(r - ra).max()))
        assert np.array_equal(e, ea), (i, members[i].level)

    for L in (2, 3, 4, 5):
        i = next(k for k, m in enumerate(members) if m.level == L)
        m = members[i]
        mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
        ro, eo = _oracle_steps(m.mesh_path, m.params, m.r0, m.r, m.seed, STEPS, mob, amp)
        r, e = got[i]
        dr = float(np.abs(r - ro).max())
        e_rel = float(np.max(np.abs(e - eo) / np.abs(eo)))
        disp = float(np.abs(ro - m.r).max())
        say("[c4] oracle member", i, "level", L, "nv", m.nv, "max |dr|", dr, "energy rel", e_rel,
            "displacement", disp)
        if precision == S.PREC_FP64:
            assert dr <= 1e-9
            assert e_rel <= 1e-10
        else:
            assert e_rel <= 1e-6
            assert dr <= 1e-5 * disp


def _serving_mix(mesh_dir):
    """Small requests (levels 2-3, 32/48 sheets) plus the largest sizes of the
    stream: the first level-6 icosphere and the first 256^2 sheet the
    generator draws, every request 500-1500 steps."""
    small = workloads.serving(n=6, rate=1e9, seed=17, levels=(2, 3), sheet_sides=(32, 48),
                              steps_range=(500, 900))
    big = workloads.serving(n=6, rate=1e9, seed=23, levels=(6, 6), sheet_sides=(256, 256),
                            steps_range=(500, 600))
    ico = next(q for q in big if q.kind == "icosphere")
    sheet = next(q for q in big if q.kind == "sheet")
    assert ico.nv == 40962 and sheet.nv == 65536
    return [small[0], ico, small[1], sheet] + small[2:]


def _serve(reqs, precision, **cfg):
    sv = S.Server(precision=precision, **cfg)
    ids = []
    for q in reqs:
        mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
        ids.append(sv.submit(q.mesh_path, q.nv, q.params, q.c0, q.steps, q.seed, mob, amp, q.r0, q.r))
    res = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    return res, st


def _request_alone(q, precision):
    s = S.Surface(q.mesh_path)
    s.set_params(precision=precision, **q.params)
    s.generatePlan()
    s.set_face_c0(np.full(s.nfaces, q.c0))
    s.setg0(q.r0)
    s.set_step(0)
    d_r = torch.tensor(q.r, device="cuda", dtype=torch.float64)
    e = torch.zeros(q.steps, device="cuda", dtype=torch.float64)
    mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
    s.bd_steps_rng_dev(d_r, q.steps, mob, amp, q.seed, e)
    torch.cuda.synchronize()
    return d_r.cpu().numpy(), e.cpu().numpy()


def test_serving_largest_requests(mesh_dir):
    reqs = _serving_mix(mesh_dir)
    say("[c5] requests", [(q.kind, q.size, q.nv, q.steps) for q in reqs])
    assert min(q.steps for q in reqs) >= 500
    res, st = _serve(reqs, S.PREC_FP64, max_active=3, chunk_steps=100)
    say("[c5] stats", st)
    assert st["done"] == len(reqs)
    for i, (q, (r, e, t)) in enumerate(zip(reqs, res)):
        ra, ea = _request_alone(q, S.PREC_FP64)
        say("[c5] request", i, q.kind, q.size, "steps", q.steps, "alone max |dr|", float(np.abs(r - ra).max()))
        assert np.array_equal(r, ra), i
        assert np.array_equal(e, ea), i
    checked = [next(i for i, q in enumerate(reqs) if q.kind == "icosphere" and q.size == L) for L in (2, 3)]
    for i in checked:
        q = reqs[i]
        mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
        ro, eo = _oracle_steps(q.mesh_path, q.params, q.r0, q.r, q.seed, q.steps, mob, amp, c0=q.c0)
        r, e, _ = res[i]
        dr = float(np.abs(r - ro).max())
        e_rel = float(np.max(np.abs(e - eo) / np.abs(eo)))
        say("[c5] oracle request", i, "level", q.size, "steps", q.steps, "max |dr|", dr, "energy rel", e_rel)
        assert dr <= 1e-9
        assert e_rel <= 1e-10
