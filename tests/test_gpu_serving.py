"""The serving loop on the GPU.

Requests of uneven size and length (icospheres of level 2-4, periodic sheets)
go through the server with continuous batching (fewer slots than requests,
chunks shorter than most requests, so requests join and leave the batch while
others run).  Bar: every request's final positions and per-step energies are
bitwise equal to the same request run alone with umcp_bd_steps_rng_dev.  A
device budget that holds only a few requests still serves them all without
exceeding it; a malformed request fails alone.
"""
import sys
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from umcp_amd import surface as S  # noqa: E402
from umcp_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def reqs(mesh_dir):
    out = workloads.serving(n=40, rate=1e9, seed=11, levels=(2, 3, 4), sheet_sides=(32, 48),
                            steps_range=(13, 160))
    return out[:8]


def alone(q, precision):
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


def serve(reqs, precision, **cfg):
    sv = S.Server(precision=precision, **cfg)
    ids = []
    for q in reqs:
        mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
        ids.append(sv.submit(q.mesh_path, q.nv, q.params, q.c0, q.steps, q.seed, mob, amp, q.r0, q.r))
    res = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    return res, st


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_MIXED])
def test_served_equals_alone(reqs, precision):
    say("[serving] precision", precision, "requests", [(q.kind, q.size, q.steps) for q in reqs])
    res, st = serve(reqs, precision, max_active=3, chunk_steps=25)
    say("   stats", st)
    assert st["done"] == len(reqs)
    assert st["peak_active"] <= 3
    for i, (q, (r, e, t)) in enumerate(zip(reqs, res)):
        ra, ea = alone(q, precision)
        say("   request", i, q.kind, q.size, "steps", q.steps, "max |dr|", float(np.abs(r - ra).max()),
            "times", t)
        assert np.array_equal(r, ra), i
        assert np.array_equal(e, ea), i
        assert t[0] <= t[1] <= t[2] <= t[3]


def test_budget_admission(reqs):
    one = S.Surface(reqs[0].mesh_path)
    say("[serving] budget")
    big = max(reqs, key=lambda q: q.nv)
    b = S.Surface(big.mesh_path)
    import ctypes as C
    planned = C.c_size_t()
    b._L.umcp_prepare(b._h)
    b._L.umcp_device_bytes(b._h, C.byref(planned), None)
    budget = 2.5 * planned.value
    res, st = serve(reqs, S.PREC_MIXED, max_active=8, chunk_steps=40, budget_bytes=budget)
    say("   budget", budget, "peak", st["peak_bytes"], "done", st["done"])
    assert st["done"] == len(reqs)
    assert st["peak_bytes"] <= budget
    del one


def test_destroy_right_after_a_failing_request(reqs, tmp_path):
    """A request whose preparation fails (missing mesh file, positions of
    another mesh) while the server is being destroyed: destroy returns
    promptly instead of leaving the worker waiting."""
    say("[serving] destroy after failing requests")
    q = reqs[0]
    other = next(x for x in reqs if x.nv != q.nv)
    mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
    for case in range(3):
        sv = S.Server(precision=S.PREC_FP64, max_active=2, chunk_steps=20)
        if case != 1:
            sv.submit(str(tmp_path / "missing.mesh"), q.nv, q.params, q.c0, 10, 1, mob, amp, q.r0, q.r)
        if case != 0:
            sv.submit(q.mesh_path, other.nv, q.params, q.c0, 10, 1, mob, amp, other.r0, other.r)
        t0 = time.time()
        sv.close()
        assert time.time() - t0 < 30.0


def test_bad_request_fails_alone(reqs):
    say("[serving] bad request")
    sv = S.Server(precision=S.PREC_MIXED, max_active=2, chunk_steps=20)
    q = reqs[0]
    mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
    good = sv.submit(q.mesh_path, q.nv, q.params, q.c0, q.steps, q.seed, mob, amp, q.r0, q.r)
    other = next(x for x in reqs if x.nv != q.nv)
    # positions of another mesh: rejected at admission
    bad = sv.submit(q.mesh_path, other.nv, q.params, q.c0, 10, 1, mob, amp, other.r0, other.r)
    with pytest.raises(S.UmcpError):
        sv.wait(bad)
    r, e, t = sv.wait(good)
    ra, _ = alone(q, S.PREC_MIXED)
    assert np.array_equal(r, ra)
    sv.close()
