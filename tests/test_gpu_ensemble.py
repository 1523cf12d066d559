"""Batched ensembles on the GPU: each member's trajectory is the same alone and
in any mix.

A mix of surfaces of uneven size (vesicles of 162, 642 and 2562 vertices with
different kc and energy forms, one with area and volume restraints, a
periodic 32x32 sheet) runs Brownian steps with device noise as one ensemble;
every member also runs alone (umcp_bd_steps_rng_dev) from the same state.
Bar: bitwise equal positions and per-step energies (the batched items run
the members' own code into the members' own partial slots, and the noise is
each member's (seed, step) stream).  The same holds for a hipGraph replay, a
persistent grid, round-robin item order, and after membership changes.  A
mixed-precision ensemble stays within the FP32 error budget of the FP64 one.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from umcp_amd import surface as S  # noqa: E402
from umcp_amd import units, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

STEPS = 12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def specs():
    out = []
    for i, (level, R, kc) in enumerate(((2, 60.0, 12.0), (3, 100.0, 25.0), (4, 200.0, 35.0))):
        w = workloads.vesicle(level, R, 0.3 * R / 100, seed=20 + i, kc_kT=kc, name="ens_ves%d" % i)
        out.append(dict(w=w, params=dict(w.params), restraints=None, seed=1000 + i))
    out[1]["params"]["mode"] = 1
    o = O.OracleSurface(out[2]["w"].mesh_path)
    A, _ = o.area(out[2]["w"].r)
    V = o.volume(out[2]["w"].r)
    out[2]["restraints"] = dict(KA_global=0.3, KV=0.02, A0=0.98 * A, V0=1.02 * V)
    ws = workloads.sheet(32, seed=6, name="ens_sheet32")
    out.append(dict(w=ws, params=dict(ws.params), restraints=None, seed=1003))
    # valence-5/7 members: a sphere and a periodic sheet (the valence-7 class)
    wd = workloads.defect_vesicle()
    out.append(dict(w=wd, params=dict(wd.params), restraints=None, seed=1004))
    wsd = workloads.defect_sheet()
    out.append(dict(w=wsd, params=dict(wsd.params), restraints=None, seed=1005))
    return out


def make_member(sp, precision=S.PREC_FP64, launch=None):
    w = sp["w"]
    g = S.Surface(w.mesh_path)
    g.set_params(precision=precision, **sp["params"])
    g.generatePlan()
    if launch:
        g.set_launch(*launch)
    if sp.get("lipids"):
        L = sp["lipids"]
        g.set_lipids(L["c0"], L["apl"], L["inner"], L["outer"], L["seed"])
    g.setg0(w.r0)
    if sp["restraints"]:
        g.set_restraints(**sp["restraints"])
    g.set_step(0)
    return g


def coeffs(sp):
    return S.bd_coefficients(1e-12, 1e-9, sp["w"].T)


def run_alone(sp, steps, precision=S.PREC_FP64, launch=None):
    g = make_member(sp, precision, launch)
    d_r = torch.tensor(sp["w"].r, device="cuda", dtype=torch.float64)
    e = torch.zeros(steps, device="cuda", dtype=torch.float64)
    mob, amp = coeffs(sp)
    g.bd_steps_rng_dev(d_r, steps, mob, amp, sp["seed"], e)
    torch.cuda.synchronize()
    return d_r.cpu(), e.cpu()


def run_ensemble(sps, steps, graph_steps=0, face_grid=0, order=0, precision=S.PREC_FP64, chunks=1,
                 launch=(4, S.LAUNCH_OCC2)):
    ens = S.Ensemble()
    ens.set_launch(launch[0], launch[1], face_grid, order)
    keep = []
    for sp in sps:
        g = make_member(sp, precision)
        d_r = torch.tensor(sp["w"].r, device="cuda", dtype=torch.float64)
        e = torch.zeros(steps, device="cuda", dtype=torch.float64)
        mob, amp = coeffs(sp)
        ens.add(g, d_r, mob, amp, sp["seed"], e)
        keep.append((g, d_r, e))
    torch.cuda.synchronize()
    per = steps // chunks
    for _ in range(chunks):
        ens.bd_steps(per, graph_steps)
    torch.cuda.synchronize()
    work = ens.work()
    out = [(d_r.cpu(), e.cpu()) for _, d_r, e in keep]
    ens.close()
    return out, work


@pytest.fixture(scope="module")
def mix():
    sps = specs()
    alone = []
    for i, sp in enumerate(sps):
        say("[ensemble] alone", i, "nv", sp["w"].nv)
        alone.append(run_alone(sp, STEPS))
    return sps, alone


def check_same(tag, got, alone):
    for i, ((r, e), (ra, ea)) in enumerate(zip(got, alone)):
        dr = float((r - ra).abs().max())
        say("   ", tag, "member", i, "max |dr|", dr, "energy equal", bool(torch.equal(e, ea)))
        assert torch.equal(r, ra), (tag, i)
        assert torch.equal(e, ea), (tag, i)


def test_ensemble_equals_alone(mix):
    sps, alone = mix
    say("[ensemble] eager, member-major")
    got, work = run_ensemble(sps, STEPS)
    say("   work items", work)
    assert work[0] > 0 and work[2] > 0
    check_same("eager", got, alone)


@pytest.mark.parametrize("graph_steps,face_grid,order", [(4, 0, 0), (0, 96, 1), (3, 40, 0)])
def test_ensemble_graph_persistent_order(mix, graph_steps, face_grid, order):
    sps, alone = mix
    say("[ensemble] graph", graph_steps, "grid", face_grid, "order", order)
    got, _ = run_ensemble(sps, STEPS, graph_steps=graph_steps, face_grid=face_grid, order=order, chunks=2)
    check_same("shape", got, alone)


def test_ensemble_membership_change(mix):
    sps, alone = mix
    say("[ensemble] membership change")
    ens = S.Ensemble()
    keep = []
    for sp in sps:
        g = make_member(sp)
        d_r = torch.tensor(sp["w"].r, device="cuda", dtype=torch.float64)
        e = torch.zeros(STEPS, device="cuda", dtype=torch.float64)
        keep.append((g, d_r, e))
    mob0, amp0 = coeffs(sps[0])
    # members 0 and 1 start; member 2 joins after 4 steps, member 0 leaves after 8
    for i in (0, 1):
        ens.add(keep[i][0], keep[i][1], *coeffs(sps[i]), sps[i]["seed"], keep[i][2])
    ens.bd_steps(4, 2)
    torch.cuda.synchronize()
    ens.add(keep[2][0], keep[2][1], *coeffs(sps[2]), sps[2]["seed"], keep[2][2])
    ens.bd_steps(4, 2)
    torch.cuda.synchronize()
    ens.remove(keep[0][0])
    ens.bd_steps(4, 2)
    ens.bd_steps(4, 0)    # member 2 catches up to STEPS
    torch.cuda.synchronize()
    assert len(ens) == 2
    del mob0, amp0
    # member 0 ran 8 steps, member 1 16, member 2 12
    r0, e0 = run_alone(sps[0], 8)
    assert torch.equal(keep[0][1].cpu(), r0)
    r2, e2 = alone[2]
    assert torch.equal(keep[2][1].cpu(), r2)
    assert torch.equal(keep[2][2].cpu(), e2)
    r1, _ = run_alone(sps[1], 16)
    assert torch.equal(keep[1][1].cpu(), r1)
    ens.close()


def test_mixed_ensemble_close_to_fp64(mix):
    sps, alone = mix
    say("[ensemble] mixed vs fp64")
    got, _ = run_ensemble(sps, STEPS, graph_steps=4, precision=S.PREC_MIXED)
    for i, ((r, e), (ra, ea)) in enumerate(zip(got, alone)):
        e_rel = float(((e - ea).abs() / ea.abs().clamp_min(1e-30)).max())
        dr = float((r - ra).abs().max())
        say("    member", i, "energy rel", e_rel, "max |dr| A", dr)
        # vesicles: the 1e-6 budget; the nearly flat periodic sheet (member 3)
        # has a small energy (small curvature), so the same absolute FP32 error
        # is ~3e-6 of it
        assert e_rel <= (1e-6 if i < 3 else 1e-5)
        assert dr <= 1e-5


@pytest.mark.parametrize("split,precision", [(2, S.PREC_FP64), (3, S.PREC_FP64), (4, S.PREC_MIXED)])
def test_cps_ensemble_equals_alone(mix, split, precision):
    """The control-point split form runs the same items alone and batched
    (persistent grid, graph replay): each member's trajectory is bitwise the
    one it has alone with that launch."""
    sps, _ = mix
    launch = (split, S.LAUNCH_CPS)
    say("[ensemble] control-point split", split, "precision", precision)
    alone = [run_alone(sp, STEPS, precision, launch) for sp in sps]
    got, _ = run_ensemble(sps, STEPS, graph_steps=4, face_grid=48, precision=precision, chunks=2, launch=launch)
    check_same("cps", got, alone)


@pytest.mark.parametrize("split", [2, 4])
def test_mfma_ensemble_equals_alone(mix, split):
    """The matrix-core mixed path runs the same items alone and batched: each
    member's trajectory is bitwise the one it has alone with that launch."""
    sps, _ = mix
    launch = (split, S.LAUNCH_MFMA)
    say("[ensemble] mixed matrix-core split", split)
    alone = [run_alone(sp, STEPS, S.PREC_MIXED, launch) for sp in sps]
    got, _ = run_ensemble(sps, STEPS, graph_steps=4, face_grid=64, precision=S.PREC_MIXED, chunks=2,
                          launch=launch)
    check_same("mfma", got, alone)


def test_ensemble_with_lipid_composition_equals_alone():
    """A member with a three-type lipid composition (lipid form, per-face
    per-type leaflet areas) batched beside a Helfrich member: both bitwise
    equal to their runs alone."""
    w = workloads.vesicle_642()
    comp = dict(c0=[-1.0 / 87.3, -1.0 / 29.4, -1.0 / 27.2], apl=[68.3, 63.4, 31.0], inner=[3.0, 2.0, 1.0],
                outer=[5.0, 0.0, 2.0], seed=11)
    sps = [dict(w=w, params=dict(w.params), restraints=None, seed=2001, lipids=comp)]
    wh = workloads.vesicle(3, 100.0, 0.3, seed=21, kc_kT=25.0, name="ens_ves1")
    sps.append(dict(w=wh, params={**wh.params, "mode": 1}, restraints=None, seed=2002))
    alone = [run_alone(sp, STEPS) for sp in sps]
    got, _ = run_ensemble(sps, STEPS)
    check_same("lipids", got, alone)
