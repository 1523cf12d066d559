"""Benchmark: Brownian dynamics on one MI355X.

Headline (value): the 40962-vertex vesicle (81920 faces: 81860 regular, 60
irregular; Helfrich form with per-face c0, global area and volume
restraints).  One step = one pass of the hot path over the resident mesh:
the restraint partials and multipliers (area of every face, volume of the
limit-point polyhedron), the fused energy + gradient of all faces, and the
per-vertex Brownian update driven by a device-resident standard-normal array
(seeded).  Timed with HIP events on the stream the work is launched on; K
steps between two device synchronisations.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line: value = FP64 face-evals/s of that loop, the roofline of
the dominant kernel (the face kernel), the FP32-mixed leg (throughput,
speedup, error against the FP64 oracle), the batched-ensemble leg (256
vesicles in one launch per kernel, hipGraph, tuned), the serving leg
(Poisson request stream, continuous batching against one-at-a-time), the
periodic 128² sheet (FP64 and mixed), the effective-mass integrators of
src/hd.C on the 642-vertex vesicle (Brownian and Langevin forms, set-up time)
and the CPU oracle timed on a bounded sample of
the headline workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "energy+grad face-evals/s & BD steps/s, FP64 vs FP32-mixed; p99 request latency"

# Algorithmic work per face-eval (energy + gradient of one face over all its
# Gauss points), in the reference algorithm's terms:
#   F = P * (2*6*ncp*3 forward + 2*5*ncp*3 scatter + 2*5*ncp*3 alpha + C_pt)
# with C_pt = 700 flop, the static operation count of the pointwise curvature
# and chain rule of src/grad.C:158-466, 470-850 as restated in the oracle.
C_PT = 700


_T0 = time.perf_counter()


def progress(msg: str):
    """One progress line on stderr per leg (stdout carries only the JSON)."""
    print("[bench %6.1f s] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


def face_flops(ncp: int, npt: int) -> int:
    return npt * (2 * 6 * ncp * 3 + 2 * 5 * ncp * 3 + 2 * 5 * ncp * 3 + C_PT)


def face_bytes(ncp: int, npt: int, nv: int, nf: int) -> float:
    """Minimal unique FP64 traffic of one face-eval: ids, g0 per point, the
    two lipid areas, and its share of reading r and writing g once."""
    return 4 * ncp + 8 * npt + 16 + 48.0 * nv / nf


PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
AMU_KG = 1.661e-27         # atomic mass unit in kg, as the reference converts masses (src/units.h)
FEAT_HELFRICH = 2         # energy-form template argument of the face kernel (mode B, kg = k_reg = 0)


def kernel_name(precision: str, split: int, flags: int, feat: int = FEAT_HELFRICH) -> str:
    """The face kernel's symbol as rocprofv3 prints it (one workgroup per item)."""
    if precision == "mixed" and flags & 4:   # matrix-core path
        return "faces_batch_mfma<%d, %d, %d, false>" % (split, 2 if flags & 2 else 1, feat)
    r = "float" if precision == "mixed" else "double"
    if flags & 8:   # control points split over the waves
        return "faces_batch_kernel<%s, %d, true, 1, %d, false, true>" % (r, split, feat)
    occ = (4 if precision == "mixed" else 2) if flags & 2 else 1
    return "faces_batch_kernel<%s, %d, true, %d, %d, false, false>" % (r, split, occ, feat)


def measured_traffic(name: str):
    """HBM bytes per launch of kernel `name` from the committed PMC summaries
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by the factor a
    plain device copy of known size reads on the same counter); None when no
    summary covers this exact kernel."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            line = line.strip()
            if not line.startswith("traffic " + name + " "):
                continue
            kv = dict(t.split("=") for t in line[len("traffic " + name) + 1:].split())
            return 1024.0 * (float(kv["fetch_kib"]) * float(kv["fetch_factor"])
                             + float(kv["write_kib"]) * float(kv["write_factor"]))
    return None


def cpu_baseline(w, c0, budget_s: float):
    """The oracle (serial C, one core) on energy+gradient evaluations of the
    same vesicle until ~budget_s of CPU work; also returns its E and gradient
    at the start configuration for the error report."""
    import numpy as np
    from oracle import oracle as O
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**w.params)
    o.set_face_c0(c0)
    o.setg0(w.r0)
    o.set_restraints(**w.extra["restraints"])
    t0 = time.perf_counter()
    n = 0
    g = np.zeros_like(w.r)
    E = g_ref = None
    while True:
        E = o.energy(w.r)
        g[:] = 0.0
        o.grad(w.r, g)
        if g_ref is None:
            g_ref = g.copy()
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    base = {"value": n * o.nfaces / dt, "unit": "face-evals/s", "cores": 1, "kind": "port",
            "sample": "%d energy+gradient evaluations of the 40962-vertex vesicle (%d faces), "
                      "serial oracle, %.1f s" % (n, o.nfaces, dt)}
    # the same evaluations with the faces split over the host's threads
    # (OpenMP, per-thread gradient buffers; at most 16, the host share of one GPU)
    nt = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))
    t0 = time.perf_counter()
    m = 0
    while True:
        o.energy_grad_threads(w.r, nt)
        m += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    dt2 = time.perf_counter() - t0
    base["threads"] = {"value": m * o.nfaces / dt2, "unit": "face-evals/s", "cores": nt, "kind": "port",
                       "sample": "%d energy+gradient evaluations, faces over %d OpenMP threads, %.1f s"
                                 % (m, nt, dt2)}
    return base, E, g_ref


def run_leg(args, w, c0, precision, d_xi, dev, torch, D=1e-12, dt=1e-9):
    """Sets up one surface, tunes the launch shape, runs W + K Brownian steps
    (diffusion D m^2/s, time step dt s)."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import tune

    s = S.Surface(w.mesh_path)
    s.set_params(precision=S.PREC_MIXED if precision == "mixed" else S.PREC_FP64, **w.params)
    s.generatePlan()
    if c0 is not None:
        s.set_face_c0(c0)
    s.setg0(w.r0)
    if w.extra and w.extra.get("restraints"):
        s.set_restraints(**w.extra["restraints"])
    classes = s.classes()
    nv, nf = s.nv, s.nfaces
    mob, amp = S.bd_coefficients(D, dt, w.T)
    K, W = args.steps, args.warmup
    stream = torch.cuda.Stream(device=dev)          # default priority
    sptr = stream.cuda_stream
    d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
    d_E = torch.zeros(1, device=dev, dtype=torch.float64)
    d_g = torch.zeros_like(d_r)
    d_en = torch.zeros(W + K, device=dev, dtype=torch.float64)
    tuned = None
    if args.split or args.flags >= 0:
        sp0, fl0 = s.get_launch()
        s.set_launch(args.split or sp0, args.flags if args.flags >= 0 else fl0)
    elif not args.no_tune:
        best, times = tune.tune(s, d_r, d_E, d_g, stream=sptr, mixed=precision == "mixed")
        tuned = {"best": list(best), "us": {"%d/%d" % k: round(v, 2) for k, v in times.items()}}
    # energy and gradient at the start configuration (for the error report)
    torch.cuda.synchronize()
    s.energy_grad_dev(d_r, d_E, d_g, stream=sptr)
    torch.cuda.synchronize()
    E0, g0 = float(d_E.cpu()[0]), d_g.cpu().numpy()

    s.bd_steps_dev(d_r, d_xi[:W], W, mob, amp, d_en[:W], stream=sptr)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the face launch of one step in 8 is bracketed by HIP events: the
    # kernel's average duration, without putting the events' cost on every step
    s.profile_begin(every=8)
    start.record(stream)
    s.bd_steps_dev(d_r, d_xi[W:W + K], K, mob, amp, d_en[W:], stream=sptr)
    end.record(stream)
    torch.cuda.synchronize()
    (face_ms, gather_ms, rs_ms), (face_n, gather_n, rs_n) = s.profile_end(3)
    ms = start.elapsed_time(end)
    energies = d_en.cpu().numpy()
    if not np.all(np.isfinite(energies)):
        raise SystemExit("non-finite energy during the timed steps (%s)" % precision)
    split, flags = s.get_launch()
    leg = {
        "face_evals_per_s": nf * K / (ms / 1e3),
        "bd_steps_per_s": K / (ms / 1e3),
        "ms_per_step": ms / K,
        "face_kernel_us": 1e3 * face_ms / face_n,
        "gather_us": 1e3 * gather_ms / max(gather_n, 1),
        "restraint_us": 1e3 * rs_ms / rs_n if rs_n else None,
        # the Brownian update itself: per vertex r read and written, the
        # gradient and the deviate read (96 B algorithmic, src/bd.C:1035); the
        # gather also streams the (face, slot) gradient records back
        "bd_update": {
            "algorithmic_bytes": 96 * nv,
            "record_bytes": 3 * (4 if precision == "mixed" else 8) * sum(c["nf"] * c["ncp"] for c in classes),
            "achieved_GB_per_s": 96 * nv / (gather_ms / max(gather_n, 1) / 1e3) / 1e9,
            "frac_of_8TB_s": 96 * nv / (gather_ms / max(gather_n, 1) / 1e3) / 8e12,
            "with_records_frac_of_8TB_s": (96 * nv + 3 * (4 if precision == "mixed" else 8) *
                                           sum(c["nf"] * c["ncp"] for c in classes)) /
                                          (gather_ms / max(gather_n, 1) / 1e3) / 8e12,
        },
        "launch": {"point_split": split, "flags": flags},
        "autotune": tuned,
        "flop_per_launch": sum(c["nf"] * face_flops(c["ncp"], c["npt"]) for c in classes),
        "bytes_per_launch": sum(c["nf"] * face_bytes(c["ncp"], c["npt"], nv, nf) for c in classes),
        "energy_first_last": [float(energies[W]), float(energies[-1])],
    }
    s.close()
    return leg, E0, g0


def ensemble_surfaces(members, precision, S):
    out = []
    for m in members:
        s = S.Surface(m.mesh_path)
        s.set_params(precision=precision, **m.params)
        s.generatePlan()
        s.setg0(m.r0)
        s.set_step(0)
        out.append(s)
    return out


def run_ensemble_leg(args, dev, torch):
    """Batched ensemble: 256 vesicles of levels 2-5 stepped together (one
    batched launch each of the face and vertex kernels per step, captured in a
    hipGraph), FP32-mixed; the launch shape picked by the ensemble tuner.
    Reports member-steps/s and face-evals/s, the same steps without the graph,
    the one-at-a-time rate measured on a sample of members, and per-member
    energy errors of the mixed run against the FP64 ensemble."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import tune, workloads

    members = workloads.ensemble(n=args.ensemble_size)
    K, W, G = args.ensemble_steps, 10, args.graph_steps
    t0 = time.perf_counter()
    surfs = ensemble_surfaces(members, S.PREC_MIXED, S)
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    d_rs = [torch.tensor(m.r, device=dev, dtype=torch.float64) for m in members]
    d_es = [torch.zeros(W + K, device=dev, dtype=torch.float64) for _ in members]
    ens = S.Ensemble()
    for m, s, d_r, d_e in zip(members, surfs, d_rs, d_es):
        mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
        ens.add(s, d_r, mob, amp, m.seed, d_e)
    torch.cuda.synchronize()
    tuned = None
    if not args.no_tune:
        progress("  ensemble launch tuner")
        best, times = tune.tune_ensemble(ens, surfs, d_rs, torch, stream, shapes=tune.ENSEMBLE_SHAPES_MIXED)
        tuned = {"best": list(best), "ms_per_step": {"%d/%d/%d/%d" % k: round(v, 4) for k, v in times.items()}}
    nf = sum(m.nfaces for m in members)
    # first W steps (graph capture included), then K timed steps
    ens.bd_steps(W, G, sptr)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    ens.bd_steps(K, G, sptr)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    energies = torch.stack(d_es).cpu().numpy()
    if not np.all(np.isfinite(energies)):
        raise SystemExit("non-finite energy in the ensemble")
    # the same launches without the graph
    Ke = min(K, 200)
    start.record(stream)
    ens.bd_steps(Ke, 0, sptr)
    end.record(stream)
    torch.cuda.synchronize()
    ms_eager = start.elapsed_time(end) / Ke
    items = ens.work()
    ens.close()
    del surfs, d_rs, d_es

    # one-at-a-time on a sample of members (device noise, no graph): the rate
    # each level reaches alone, weighted by the ensemble's composition
    rate_alone = {}
    Ks = 100
    for L in sorted({m.level for m in members}):
        m = next(x for x in members if x.level == L)
        s = ensemble_surfaces([m], S.PREC_MIXED, S)[0]
        d_r = torch.tensor(m.r, device=dev, dtype=torch.float64)
        mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
        s.bd_steps_rng_dev(d_r, 10, mob, amp, m.seed, None, sptr)
        torch.cuda.synchronize()
        start.record(stream)
        s.bd_steps_rng_dev(d_r, Ks, mob, amp, m.seed, None, sptr)
        end.record(stream)
        torch.cuda.synchronize()
        rate_alone[L] = start.elapsed_time(end) / Ks      # ms per step alone
        s.close()
    seq_ms_per_step = sum(rate_alone[m.level] for m in members)

    # mixed against FP64, per member, over the first Kc steps from the start
    progress("  ensemble: mixed against FP64")
    Kc = 20
    res = {}
    for prec in (S.PREC_FP64, S.PREC_MIXED):
        ss = ensemble_surfaces(members, prec, S)
        rr = [torch.tensor(m.r, device=dev, dtype=torch.float64) for m in members]
        ee = [torch.zeros(Kc, device=dev, dtype=torch.float64) for _ in members]
        en = S.Ensemble()
        for m, s, d_r, d_e in zip(members, ss, rr, ee):
            mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
            en.add(s, d_r, mob, amp, m.seed, d_e)
        en.bd_steps(Kc, Kc, sptr)
        torch.cuda.synchronize()
        res[prec] = ([x.cpu().numpy() for x in rr], torch.stack(ee).cpu().numpy())
        en.close()
        del ss
    r64, e64 = res[S.PREC_FP64]
    r32, e32 = res[S.PREC_MIXED]
    e_rel = np.abs(e32 / e64 - 1).max(axis=1)
    dr = np.array([np.abs(a - b).max() for a, b in zip(r32, r64)])

    levels = {str(L): sum(1 for m in members if m.level == L) for L in sorted({m.level for m in members})}
    cpu = None
    if not args.no_cpu_baseline:
        progress("  ensemble: CPU baseline")
        # the serial oracle's Brownian step, member after member in the
        # ensemble's order until the budget is spent; the ensemble step's CPU
        # time is projected from the measured face-evals/s
        from oracle import oracle as O
        rng = np.random.default_rng(4)
        t_cpu, n_m, n_f = 0.0, 0, 0
        for m in members:
            o = O.OracleSurface(m.mesh_path)
            o.set_params(**m.params)
            o.setg0(m.r0)
            mob, amp = S.bd_coefficients(1e-12, m.dt, m.T)
            xi = rng.standard_normal((1, 3 * o.nv))
            t0 = time.perf_counter()
            o.bd_steps(m.r, xi, mob, amp)
            t_cpu += time.perf_counter() - t0
            n_m, n_f = n_m + 1, n_f + o.nfaces
            del o
            if t_cpu > args.cpu_budget / 2:
                break
        rate = n_f / t_cpu
        cpu = {"value": rate, "unit": "face-evals/s", "cores": 1, "kind": "port",
               "ensemble_ms_per_step": 1e3 * nf / rate,
               "sample": "one Brownian step of each of the first %d members (%d faces), serial oracle, %.1f s"
                         % (n_m, n_f, t_cpu)}
    return {
        "workload": "ensemble_%d_levels2-5" % len(members),
        "members": len(members), "levels": levels, "faces": nf, "vertices": sum(m.nv for m in members),
        "precision": "mixed", "steps": K, "graph_steps": G,
        "member_steps_per_s": len(members) * K / (ms / 1e3),
        "face_evals_per_s": nf * K / (ms / 1e3),
        "ms_per_step": ms / K,
        "ms_per_step_no_graph": ms_eager,
        "launch": tuned["best"] if tuned else [4, 2, 0, 0],
        "autotune": tuned,
        "work_items": {"face": items[0], "gather": items[1], "restraint": items[2]},
        "one_at_a_time_ms_per_step_by_level": {str(k): round(v, 4) for k, v in rate_alone.items()},
        "speedup_vs_one_at_a_time": seq_ms_per_step / (ms / K),
        "vs_fp64": {"steps": Kc, "energy_rel_max": float(e_rel.max()), "energy_rel_median": float(np.median(e_rel)),
                    "max_dr_A": float(dr.max())},
        "setup_s": setup_s,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (cpu["ensemble_ms_per_step"] / (ms / K)) if cpu else None,
    }


def run_c1_leg(args, dev, torch):
    """The 642-vertex vesicle (kc 20 kT, c0 0, KA 0.215 kT/A^2): one energy +
    gradient evaluation, then overdamped Brownian steps at dt 1e-9 s and
    D 1e-12 m^2/s with device N(0,1) noise, FP64 and FP32-mixed; beside them
    the oracle's serial and threaded steps (host normals) on a bounded sample
    of the same steps."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import workloads
    w = workloads.vesicle_642()
    mob, amp = S.bd_coefficients(1e-12, 1e-9, w.T)
    K = args.c1_steps
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    out = {"workload": "vesicle_642_bd", "nv": w.nv, "faces": 1280, "dt_s": 1e-9, "D_m2_per_s": 1e-12,
           "steps": K}
    ends = {}
    for prec, code in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        s = S.Surface(w.mesh_path)
        s.set_params(precision=code, **w.params)
        s.generatePlan()
        s.setg0(w.r0)
        d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
        d_E = torch.zeros(1, device=dev, dtype=torch.float64)
        d_g = torch.zeros_like(d_r)
        d_en = torch.zeros(K, device=dev, dtype=torch.float64)
        s.energy_grad_dev(d_r, d_E, d_g, stream=sptr)     # warm the kernels
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        s.energy_grad_dev(d_r, d_E, d_g, stream=sptr)
        end.record(stream)
        torch.cuda.synchronize()
        eval_ms = start.elapsed_time(end)
        s.set_step(0)
        start.record(stream)
        s.bd_steps_rng_dev(d_r, K, mob, amp, 21, d_en, sptr)
        end.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / K
        e = d_en.cpu().numpy()
        ends[prec] = d_r.cpu().numpy()
        out[prec] = {"energy_grad_ms": eval_ms, "ms_per_step": ms, "steps_per_s": 1e3 / ms,
                     "face_evals_per_s": s.nfaces * 1e3 / ms, "energy_first_last": [float(e[0]), float(e[-1])],
                     "finite": bool(np.isfinite(e).all())}
        s.close()
    out["mixed"]["max_dr_vs_fp64_A"] = float(np.abs(ends["mixed"] - ends["fp64"]).max())
    out["mixed"]["displacement_A"] = float(np.abs(ends["fp64"] - w.r).max())
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        o = O.OracleSurface(w.mesh_path)
        o.set_params(**w.params)
        o.setg0(w.r0)
        nt = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))
        chunk = 25
        xi = np.random.default_rng(2).standard_normal((chunk, 3 * o.nv))
        base = {}
        for name, cores in (("serial", 1), ("threads", nt)):
            r = w.r.copy()
            n, t0 = 0, time.perf_counter()
            while n < K and time.perf_counter() - t0 < args.cpu_budget / 2:
                r = o.bd_steps(r, xi, mob, amp) if cores == 1 else o.bd_steps_threads(r, xi, mob, amp, cores)
                n += chunk
            dt = time.perf_counter() - t0
            base[name] = {"value": n / dt, "unit": "BD steps/s", "cores": cores, "kind": "port",
                          "sample": "%d of the %d Brownian steps, %s oracle, %.1f s"
                                    % (n, K, "serial" if cores == 1 else "%d-thread OpenMP" % cores, dt)}
        out["cpu_baseline"] = dict(base["serial"], threads=base["threads"])
        out["fp64"]["speedup_vs_cpu"] = out["fp64"]["steps_per_s"] / base["serial"]["value"]
        out["fp64"]["speedup_vs_cpu_threads"] = out["fp64"]["steps_per_s"] / base["threads"]["value"]
    return out


def run_c2_leg(args, dev, torch):
    """The periodic 128 x 128 hex sheet (16384 vertices, 32768 regular faces,
    box 1000 x 866 A, Fourier heights of kc 20 kT at 310 K, lipid form): the
    Langevin integrator of src/hd.C with the effective mass (OLD_LANGEVIN
    friction -Delta gamma Mi pp / alpha^2, src/hd.C:1546-1557) at friction
    gamma = 1e-10 kg/s, i.e. 1e-10 kg/s in amu per AKMA time unit (the
    reference's unit of gamma_langevin: mass over time), then 10000 steps in
    FP64 and in FP32-mixed.  The stated dt = 1e-9 s is tried first, then
    shorter steps: each trial runs 200 steps from the start state and counts
    as stable when every energy stays finite and within 100x the start
    energy plus 3 nv kT; the timed runs use the largest stable dt.  The CPU
    comparator is the serial oracle's Langevin step with the inverse from the
    GPU set-up (the reference's dense O(nv^3) serial inversion is not timed)."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import units, workloads
    w = workloads.sheet_128()
    gamma = 1e-10 / AMU_KG / units.AKMA_TIME      # amu per AKMA time unit
    kT = units.kT(w.T)
    K = args.c2_steps
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    out = {"workload": "sheet_128_periodic_langevin", "nv": w.nv, "faces": 2 * w.nv, "box_A": [1000.0, 866.0],
           "gamma_kg_per_s": 1e-10, "gamma_langevin": gamma, "T_K": w.T, "dt_requested_s": 1e-9, "steps": K}
    legs, trials, dt = {}, [], None
    for prec, code in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        s = S.Surface(w.mesh_path)
        s.set_params(precision=code, **w.params)
        s.generatePlan()
        s.setg0(w.r0)
        t0 = time.perf_counter()
        nnz, m = s.effm_build()
        setup_s = time.perf_counter() - t0
        if dt is None:
            e_cap = None
            for cand in (1e-9, 3e-10, 1e-10, 3e-11, 1e-11, 3e-12, 1e-12):
                d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
                d_pp = torch.zeros(3 * s.nv, device=dev, dtype=torch.float64)
                d_e = torch.zeros(200, device=dev, dtype=torch.float64)
                s.set_step(0)
                s.effm_steps_dev(d_r, d_pp, 200, S.EFFM_LD, cand, kT, gamma, seed=11, d_energies=d_e, stream=sptr)
                torch.cuda.synchronize()
                e = d_e.cpu().numpy()
                if e_cap is None:
                    e_cap = 100.0 * (abs(float(e[0])) + 3 * s.nv * kT)
                ok = bool(np.isfinite(e).all() and torch.isfinite(d_r).all() and np.abs(e).max() < e_cap)
                trials.append({"dt_s": cand, "stable": ok, "energy_last": float(e[-1])})
                if ok:
                    dt = cand
                    break
            if dt is None:
                raise SystemExit("sheet Langevin: no stable time step down to 1e-12 s")
        d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
        d_pp = torch.zeros(3 * s.nv, device=dev, dtype=torch.float64)
        d_e = torch.zeros(K, device=dev, dtype=torch.float64)
        s.set_step(0)
        s.effm_steps_dev(d_r, d_pp, 3, S.EFFM_LD, dt, kT, gamma, seed=12, stream=sptr)   # warm-up
        torch.cuda.synchronize()   # the reset below runs on torch's stream
        d_r.copy_(torch.tensor(w.r, device=dev, dtype=torch.float64))
        d_pp.zero_()
        s.set_step(0)
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        s.effm_steps_dev(d_r, d_pp, K, S.EFFM_LD, dt, kT, gamma, seed=12, d_energies=d_e, stream=sptr)
        end.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / K
        e = d_e.cpu().numpy()
        legs[prec] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "face_evals_per_s": 2 * w.nv * 1e3 / ms,
                      "setup_s": setup_s, "inverse_mass_kept": nnz, "point_mass_amu": m,
                      "energy_first_last": [float(e[0]), float(e[-1])], "finite": bool(np.isfinite(e).all()),
                      "r_end": d_r.cpu().numpy()}
        if prec == "fp64" and not args.no_cpu_baseline:
            Mi = s.effm_get()
        if prec == "fp64":
            # the scalable form on the same steps (M sparse, Mi pp by the
            # Chebyshev solve)
            legs["fp64_sparse_mass"] = _effm_sparse_run(s, w, dt, K, kT, gamma, dev, torch, stream,
                                                        r_ref=legs["fp64"]["r_end"], seed=12)
        s.close()
        del d_r, d_pp, d_e
    out["dt_trials"] = trials
    out["dt_s"] = dt
    r64, r32 = legs["fp64"].pop("r_end"), legs["mixed"].pop("r_end")
    legs["mixed"]["max_dr_vs_fp64_A"] = float(np.abs(r32 - r64).max())
    legs["mixed"]["energy_rel_vs_fp64_last"] = abs(legs["mixed"]["energy_first_last"][1] /
                                                   legs["fp64"]["energy_first_last"][1] - 1)
    legs["fp64"]["displacement_A"] = float(np.abs(r64 - w.r).max())
    out["fp64"], out["mixed"] = legs["fp64"], legs["mixed"]
    out["fp64_sparse_mass"] = legs["fp64_sparse_mass"]
    out["mixed_speedup_vs_fp64"] = legs["fp64"]["ms_per_step"] / legs["mixed"]["ms_per_step"]
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        o = O.OracleSurface(w.mesh_path)
        o.set_params(**w.params)
        o.setg0(w.r0)
        n = 4
        xi = np.random.default_rng(3).standard_normal((n, 3 * o.nv))
        t0 = time.perf_counter()
        o.effm_steps(w.r, np.zeros(3 * o.nv), xi, S.EFFM_LD, dt, kT, gamma, Mi)
        t = time.perf_counter() - t0
        del Mi
        out["cpu_baseline"] = {"value": n / t, "unit": "Langevin steps/s", "cores": 1, "kind": "port",
                               "sample": "%d Langevin steps of the sheet, serial oracle, inverse effective mass "
                                         "from the GPU set-up, %.1f s" % (n, t)}
        out["fp64"]["speedup_vs_cpu"] = out["fp64"]["steps_per_s"] / out["cpu_baseline"]["value"]
    return out


def run_points_leg(args, dev, torch):
    """Points on the limit surface (evaluateRNRM, pointGradient, the point
    geometry, penergy's particle terms) batched on the 40962-vertex vesicle:
    n random points (faces uniform, (u, v) uniform in the triangle), wall
    time of the host-array calls (staging, the batch's vertex incidence,
    kernels, results), and the serial oracle on a bounded sample."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import workloads
    w = workloads.vesicle_40962()
    s = S.Surface(w.mesh_path)
    s.set_params(precision=S.PREC_FP64, **w.params)
    s.generatePlan()
    n = args.points_n
    rng = np.random.default_rng(21)
    faces = rng.integers(0, s.nfaces, n).astype(np.int32)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    uv = np.stack([u, v], 1)
    a, b = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    r = w.r.copy()
    out = {"workload": "vesicle_40962_points", "nv": s.nv, "faces": s.nfaces, "points": n, "precision": "fp64"}
    s.point_eval(r, faces[:1000], uv[:1000])
    s.point_grad(r, faces[:1000], uv[:1000], a[:1000], b[:1000])
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        R, N = s.point_eval(r, faces, uv)
    te = (time.perf_counter() - t0) / reps
    g = np.zeros_like(r)
    t0 = time.perf_counter()
    for _ in range(reps):
        g[:] = 0.0
        s.point_grad(r, faces, uv, a, b, on_surface=True, kT=0.6, g=g)
    tg = (time.perf_counter() - t0) / reps
    out["eval_s"] = te
    out["eval_points_per_s"] = n / te
    # the device-resident evaluation: inputs and outputs in HBM, HIP events
    # on the stream it runs on (resolve, partition by kind, point kernels)
    stream = torch.cuda.Stream(device=dev)
    d_r = torch.tensor(r, device=dev, dtype=torch.float64)
    d_f = torch.tensor(faces, device=dev, dtype=torch.int32)
    d_uv = torch.tensor(uv, device=dev, dtype=torch.float64)
    d_R = torch.empty((n, 3), device=dev, dtype=torch.float64)
    d_N = torch.empty_like(d_R)
    s.point_eval_dev(d_r, d_f, d_uv, d_R, d_N, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    s.point_status()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(10):
        s.point_eval_dev(d_r, d_f, d_uv, d_R, d_N, stream=stream.cuda_stream)
    end.record(stream)
    torch.cuda.synchronize()
    out["eval_dev_s"] = start.elapsed_time(end) / 10 * 1e-3
    out["eval_dev_points_per_s"] = n / out["eval_dev_s"]
    out["eval_dev_equals_host"] = bool(np.array_equal(d_R.cpu().numpy(), R) and np.array_equal(d_N.cpu().numpy(), N))
    del d_r, d_f, d_uv, d_R, d_N
    out["grad_s"] = tg
    out["grad_points_per_s"] = n / tg
    # the device-resident gradient (inputs and the accumulating outputs in HBM)
    stream = torch.cuda.Stream(device=dev)
    t = lambda x, dt=torch.float64: torch.tensor(x, device=dev, dtype=dt)  # noqa: E731
    d_r, d_f, d_uv, d_a, d_b = t(r), t(faces, torch.int32), t(uv), t(a), t(b)
    d_g = torch.zeros(len(r), device=dev, dtype=torch.float64)
    d_gp = torch.zeros((n, 2), device=dev, dtype=torch.float64)
    s.point_grad_dev(d_r, d_f, d_uv, d_a, d_b, d_g, d_gp, on_surface=True, kT=0.6, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    s.point_status()
    out["grad_dev_equals_host"] = bool(np.array_equal(d_g.cpu().numpy(), g))
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(10):
        s.point_grad_dev(d_r, d_f, d_uv, d_a, d_b, d_g, d_gp, on_surface=True, kT=0.6, stream=stream.cuda_stream)
    end.record(stream)
    torch.cuda.synchronize()
    out["grad_dev_s"] = start.elapsed_time(end) / 10 * 1e-3
    out["grad_dev_points_per_s"] = n / out["grad_dev_s"]
    del d_r, d_f, d_uv, d_a, d_b, d_g, d_gp
    # the geometry query (ru, rv, g, c)
    s.point_geometry(r, faces[:1000], uv[:1000])
    t0 = time.perf_counter()
    for _ in range(reps):
        Gm = s.point_geometry(r, faces, uv)
    out["geometry_s"] = (time.perf_counter() - t0) / reps
    out["geometry_points_per_s"] = n / out["geometry_s"]
    # penergy's terms for n particles (c0 and area per particle), including
    # the face densities of every face
    s.setg0(w.r0)
    pc0 = rng.uniform(-0.03, 0.03, n)
    pa = rng.uniform(20.0, 80.0, n)
    s.particle_energy(r, faces[:1000], uv[:1000], pc0[:1000], pa[:1000])
    t0 = time.perf_counter()
    for _ in range(reps):
        Ep = s.particle_energy(r, faces, uv, pc0, pa)
    out["particle_energy_s"] = (time.perf_counter() - t0) / reps
    out["particle_energy_per_s"] = n / out["particle_energy_s"]
    out["finite"] = bool(np.isfinite(R).all() and np.isfinite(g).all() and np.isfinite(Gm).all()
                         and np.isfinite(Ep).all())
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        o = O.OracleSurface(w.mesh_path)
        o.set_params(**w.params)
        m = min(n, 20000)
        t0 = time.perf_counter()
        o.point_grad(r, faces[:m], uv[:m], a[:m], b[:m], on_surface=1, kT=0.6)
        tc = time.perf_counter() - t0
        out["cpu_oracle_grad_points_per_s"] = m / tc
        out["cpu_sample_points"] = m
    s.close()
    return out


def _effm_sparse_run(s, w, dt, K, kT, gamma, dev, torch, stream, r_ref=None, seed=7):
    """K device-noise Langevin steps (after 3 warm-up steps, from the start
    positions and step 0) with the scalable effective mass
    (umcp_effm_build_sparse) on surface s; set-up time, ms per step, the
    solve's length and, given the dense form's end positions for the same
    steps, the largest difference to them."""
    import numpy as np
    from umcp_amd import surface as S
    sptr = stream.cuda_stream
    t0 = time.perf_counter()
    nnz, m, iters = s.effm_build_sparse()
    setup = time.perf_counter() - t0
    d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
    d_pp = torch.zeros(3 * s.nv, device=dev, dtype=torch.float64)
    s.set_step(0)
    s.effm_steps_dev(d_r, d_pp, 3, S.EFFM_LD, dt, kT, gamma, seed=seed, stream=sptr)   # warm-up
    torch.cuda.synchronize()   # the reset below runs on torch's stream
    d_r.copy_(torch.tensor(w.r, device=dev, dtype=torch.float64))
    d_pp.zero_()
    s.set_step(0)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    s.effm_steps_dev(d_r, d_pp, K, S.EFFM_LD, dt, kT, gamma, seed=seed, stream=sptr)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / K
    out = {"nv": s.nv, "mass_nnz": nnz, "solve_iterations": iters, "setup_s": setup, "dt_s": dt,
           "langevin_steps": K, "langevin_ms_per_step": ms, "finite": bool(torch.isfinite(d_r).all()),
           # M's entries (8 + 4 bytes) and per row x, r read and written, d
           # read and written, 1/M_pp: 152 bytes
           "bytes_per_solve_iteration": nnz * 12 + s.nv * 152}
    if r_ref is not None:
        out["max_dr_vs_dense_A"] = float(np.abs(d_r.cpu().numpy() - r_ref).max())
        out["displacement_A"] = float(np.abs(r_ref - w.r).max())
    return out


def run_effm_leg(args, dev, torch):
    """Momentum-space integrators of src/hd.C with the effective-mass matrix
    on the 642-vertex vesicle (the size the reference's O(nv^3) set-up
    serves): set-up time (mass matrix, Cholesky inverse, truncation, upload),
    then K device-noise steps of the Brownian and of the Langevin form, and
    the serial oracle on a bounded sample of the same steps."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import units, workloads
    w = workloads.vesicle_642()
    s = S.Surface(w.mesh_path)
    s.set_params(precision=S.PREC_FP64, **w.params)
    s.generatePlan()
    s.setg0(w.r0)
    t0 = time.perf_counter()
    nnz, m = s.effm_build()
    build_s = time.perf_counter() - t0
    dt, gamma, kT = 1e-11, 0.001, units.kT(298.0)
    K = args.effm_steps
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    out = {"workload": "vesicle_642_effm", "nv": s.nv, "faces": s.nfaces, "precision": "fp64",
           "inverse_mass_nnz": nnz, "point_mass_amu": m, "setup_s": build_s, "dt_s": dt,
           "gamma_langevin": gamma, "steps": K}
    for name, mode in (("brownian", S.EFFM_BD), ("langevin", S.EFFM_LD)):
        d_r = torch.tensor(w.r, device=dev, dtype=torch.float64)
        d_pp = torch.zeros(3 * s.nv, device=dev, dtype=torch.float64)
        s.set_step(0)
        s.effm_steps_dev(d_r, d_pp, 5, mode, dt, kT, gamma, seed=7, stream=sptr)
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        s.effm_steps_dev(d_r, d_pp, K, mode, dt, kT, gamma, seed=7, stream=sptr)
        end.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / K
        out[name] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "face_evals_per_s": s.nfaces * 1e3 / ms,
                     "max_displacement_A": float((d_r - torch.tensor(w.r, device=dev)).abs().max())}
    # the same integrator on the 40962-vertex vesicle: a set-up the
    # reference's dense CPU inverse cannot reach (13.4 GB inverse in HBM,
    # kept as CSR after the truncation)
    wl = workloads.vesicle_40962()
    sl = S.Surface(wl.mesh_path)
    sl.set_params(precision=S.PREC_FP64, **wl.params)
    sl.generatePlan()
    sl.setg0(wl.r0)
    t0 = time.perf_counter()
    nnz_l, m_l = sl.effm_build()
    build_l = time.perf_counter() - t0
    d_r = torch.tensor(wl.r, device=dev, dtype=torch.float64)
    d_pp = torch.zeros(3 * sl.nv, device=dev, dtype=torch.float64)
    sl.set_step(0)
    dt_l, K_l = 1e-12, max(20, K // 5)
    sl.effm_steps_dev(d_r, d_pp, 3, S.EFFM_LD, dt_l, kT, gamma, seed=7, stream=sptr)
    torch.cuda.synchronize()   # the reset below runs on torch's stream
    d_r.copy_(torch.tensor(wl.r, device=dev, dtype=torch.float64))
    d_pp.zero_()
    sl.set_step(0)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    sl.effm_steps_dev(d_r, d_pp, K_l, S.EFFM_LD, dt_l, kT, gamma, seed=7, stream=sptr)
    end.record(stream)
    torch.cuda.synchronize()
    ms_l = start.elapsed_time(end) / K_l
    bytes_l = 10 * nnz_l if 5 * nnz_l < 4 * sl.nv ** 2 else 8 * sl.nv ** 2
    out["vesicle_40962"] = {"nv": sl.nv, "inverse_mass_kept": nnz_l, "kept_fraction": nnz_l / sl.nv ** 2,
                            "setup_s": build_l, "dt_s": dt_l, "langevin_steps": K_l,
                            "langevin_ms_per_step": ms_l, "inverse_bytes_per_product": bytes_l,
                            "product_GB_per_s_upper_bound": bytes_l / (ms_l * 1e-3) / 1e9,
                            "finite": bool(torch.isfinite(d_r).all())}
    r_dense = d_r.cpu().numpy()
    # the scalable form on the same surface and steps: M sparse, Mi pp by the
    # Chebyshev solve
    out["vesicle_40962"]["sparse_mass"] = _effm_sparse_run(sl, wl, dt_l, K_l, kT, gamma, dev, torch, stream,
                                                           r_ref=r_dense)
    sl.close()
    del d_r, d_pp
    # beyond the dense form's limit: 163842 vertices (the dense inverse would
    # be 215 GB)
    wx = workloads.vesicle(7, 1600.0, 0.5, seed=12, name="effm_ves163842")
    sx = S.Surface(wx.mesh_path)
    sx.set_params(precision=S.PREC_FP64, **wx.params)
    sx.generatePlan()
    sx.setg0(wx.r0)
    out["vesicle_163842"] = _effm_sparse_run(sx, wx, dt_l, max(10, K_l // 2), kT, gamma, dev, torch, stream)
    sx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        o = O.OracleSurface(w.mesh_path)
        o.set_params(**w.params)
        o.setg0(w.r0)
        t0 = time.perf_counter()
        Mi = o.effm_inverse()
        t_setup = time.perf_counter() - t0
        n = 20
        xi = np.random.default_rng(1).standard_normal((n, 3 * o.nv))
        t0 = time.perf_counter()
        o.effm_steps(w.r, np.zeros(3 * o.nv), xi, S.EFFM_LD, dt, kT, gamma, Mi)
        t_steps = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / t_steps, "unit": "Langevin steps/s", "cores": 1, "kind": "port",
                               "setup_s": t_setup,
                               "sample": "%d Langevin steps of the 642-vertex vesicle, serial oracle" % n}
        out["langevin"]["speedup_vs_cpu"] = out["langevin"]["steps_per_s"] / out["cpu_baseline"]["value"]
    return out


def run_serving_leg(args, dev, torch):
    """Serving stream: Poisson arrivals of Brownian-dynamics requests
    (icospheres of level 2-6 or periodic sheets of 32^2..256^2 vertices,
    log-uniform steps in [100, 10000], per-request kc and c0), FP32-mixed,
    through the serving loop with continuous batching.  The stream of
    --serve-requests at --serve-rate gives the latency percentiles (submit to
    done) and aggregate throughput over the makespan; its first
    --serve-alone requests also run through the same loop one at a time (the
    comparator: face-evals/s of the two runs, since the subset's mix differs
    from the whole stream's; and every request's result equal in both).  A
    latency-versus-load curve repeats the stream's first --serve-curve
    requests at 10, 20, 50 and 200 requests/s, and a stream of small requests
    (levels 2-3, sheets of 32^2..64^2, 100-1000 steps) offered at 2000/s runs
    both ways."""
    import numpy as np
    from umcp_amd import surface as S
    from umcp_amd import workloads

    def run(reqs, slots):
        sv = S.Server(max_active=slots, chunk_steps=args.serve_chunk, precision=S.PREC_MIXED)
        # the mesh library is loaded before the stream starts: one zero-step
        # request per distinct mesh file parses and plans its template
        first = {}
        for q in reqs:
            first.setdefault(q.mesh_path, q)
        for q in first.values():
            sv.wait(sv.submit(q.mesh_path, q.nv, q.params, q.c0, 0, q.seed, 0.0, 0.0, q.r0, q.r))
        st0 = sv.stats()
        ids = []
        t0 = time.perf_counter()
        for q in reqs:
            delay = t0 + q.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            mob, amp = S.bd_coefficients(1e-12, q.dt, q.T)
            ids.append(sv.submit(q.mesh_path, q.nv, q.params, q.c0, q.steps, q.seed, mob, amp, q.r0, q.r))
        res = [sv.wait(i) for i in ids]
        st = sv.stats()
        sv.close()
        t_sub = np.array([t[0] for _, _, t in res])
        t_done = np.array([t[3] for _, _, t in res])
        lat = t_done - t_sub
        makespan = t_done.max() - t_sub.min()
        face_steps = float(sum(q.nfaces * q.steps for q in reqs))
        return {
            "slots": slots,
            "requests": len(reqs),
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "mean": float(lat.mean()), "max": float(lat.max()), "count": len(lat),
                          "above_p99": int((lat > np.percentile(lat, 99)).sum())},
            "queue_wait_s_p50": float(np.percentile(np.array([t[1] - t[0] for _, _, t in res]), 50)),
            "makespan_s": float(makespan),
            "requests_per_s": len(reqs) / makespan,
            "face_evals_per_s": face_steps / makespan,
            "device_busy_fraction": (st["busy_s"] - st0["busy_s"]) / makespan,
            "worker_s": {k: st[k + "_s"] - st0[k + "_s"] for k in ("busy", "build", "retire", "wait")},
            "setup_s_total": st["setup_s"] - st0["setup_s"],
            "chunks": int(st["chunks"]), "peak_active": int(st["peak_active"]),
            "peak_reserved_GB": st["peak_bytes"] / 1e9,
        }, res

    def same(ra, rb):
        return {"max_abs_dr_A": max(float(np.abs(a[0] - b[0]).max()) for a, b in zip(ra, rb)),
                "energies_bitwise_equal": bool(all(np.array_equal(a[1], b[1]) for a, b in zip(ra, rb)))}

    reqs = workloads.serving(n=args.serve_requests, rate=args.serve_rate)
    progress("  %d requests at %g/s, continuous batching" % (len(reqs), args.serve_rate))
    batched, res_b = run(reqs, args.serve_slots)
    sub = reqs[:args.serve_alone]
    progress("  the first %d one at a time" % len(sub))
    single, res_1 = run(sub, 1)
    curve = []
    for rate in (10.0, 20.0, 50.0, 200.0):
        rq = workloads.serving(n=args.serve_curve, rate=rate)
        progress("  load curve: %d requests at %g/s" % (len(rq), rate))
        c, _ = run(rq, args.serve_slots)
        curve.append({"offered_requests_per_s": rate, "requests": c["requests"],
                      "p50_s": c["latency_s"]["p50"], "p99_s": c["latency_s"]["p99"],
                      "max_s": c["latency_s"]["max"], "requests_per_s": c["requests_per_s"],
                      "face_evals_per_s": c["face_evals_per_s"], "queue_wait_s_p50": c["queue_wait_s_p50"],
                      "peak_active": c["peak_active"]})
    # offered faster than either way serves them, so both runs measure capacity
    small = workloads.serving(n=args.serve_small, rate=2000.0, seed=9, levels=(2, 3),
                              sheet_sides=(32, 64), steps_range=(100, 1000))
    progress("  %d small requests, both ways" % len(small))
    sm_b, rsb = run(small, args.serve_slots)
    sm_1, rs1 = run(small, 1)
    return {
        "workload": "serving_%d_requests_poisson_%g_per_s" % (len(reqs), args.serve_rate),
        "precision": "mixed",
        "requests": len(reqs),
        "icospheres": sum(1 for q in reqs if q.kind == "icosphere"),
        "sheets": sum(1 for q in reqs if q.kind == "sheet"),
        "face_steps": float(sum(q.nfaces * q.steps for q in reqs)),
        "continuous_batching": batched,
        "one_at_a_time": dict(single, note="the stream's first %d requests" % len(sub)),
        "throughput_ratio": batched["face_evals_per_s"] / single["face_evals_per_s"],
        "throughput_ratio_unit": "face-evals/s of the stream over face-evals/s of its first requests one at a time",
        "same_result_alone_and_in_mix": same(res_b[:len(sub)], res_1),
        "curve": curve,
        "small_requests": {"requests": len(small), "continuous_batching": sm_b, "one_at_a_time": sm_1,
                           "throughput_ratio": sm_b["requests_per_s"] / sm_1["requests_per_s"],
                           "same_result_alone_and_in_mix": same(rsb, rs1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--split", type=int, default=0, help="fix the point split (skip tuning)")
    ap.add_argument("--flags", type=int, default=-1, help="fix the launch flags (skip tuning)")
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default launch shape")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-mixed leg")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-size", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--graph-steps", type=int, default=50)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--no-effm", action="store_true", help="skip the effective-mass integrator leg")
    ap.add_argument("--no-sheet", action="store_true", help="skip the periodic-sheet Langevin leg")
    ap.add_argument("--c2-steps", type=int, default=10000)
    ap.add_argument("--no-c1", action="store_true", help="skip the 642-vertex Brownian leg")
    ap.add_argument("--c1-steps", type=int, default=1000)
    ap.add_argument("--effm-steps", type=int, default=500)
    ap.add_argument("--no-points", action="store_true", help="skip the surface-points leg")
    ap.add_argument("--points-n", type=int, default=1000000)
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-alone", type=int, default=128, help="requests of the stream also run one at a time")
    ap.add_argument("--serve-curve", type=int, default=128, help="requests per point of the load curve")
    ap.add_argument("--serve-small", type=int, default=256, help="requests of the small-request stream")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-chunk", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs one mesh per GPU and has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch  # device memory and streams; its HIP runtime is shared with the library

    from umcp_amd import workloads

    dev = torch.device("cuda:0")
    w = workloads.vesicle_40962()
    c0 = workloads.face_c0_uniform(81920, w.extra["c0_seed"])
    K, W = args.steps, args.warmup
    gen = torch.Generator(device=dev)
    gen.manual_seed(12345)
    d_xi = torch.randn((W + K, 3 * w.nv), generator=gen, device=dev, dtype=torch.float64)

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(n, device=dev, dtype=torch.float64)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    progress("40962-vertex vesicle, FP64: tuning and %d Brownian steps" % K)
    leg64, E64, g64 = run_leg(args, w, c0, "fp64", d_xi, dev, torch)
    face_s = leg64["face_kernel_us"] / 1e6
    achieved = leg64["flop_per_launch"] / face_s / 1e12
    fl = leg64["launch"]["flags"]
    kname = kernel_name("fp64", leg64["launch"]["point_split"], fl)
    out = {
        "metric": METRIC,
        "value": leg64["face_evals_per_s"],
        "unit": "face-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": leg64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded icosphere, seeded device-resident N(0,1) noise)",
        "config": {"workload": "vesicle_40962_bd", "nv": w.nv, "faces": 81920,
                   "faces_regular": 81860, "faces_irregular": 60,
                   "energy_form": "helfrich per-face c0", "restraints": w.extra["restraints"],
                   "precision": "fp64",
                   "D_m2_per_s": 1e-12, "dt_s": 1e-9},
        "bd_steps_per_s": leg64["bd_steps_per_s"],
        "fp64": leg64,
        "roofline": {"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": measured_traffic(kname),
                     "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "kernel": kname,
                     "avg_launch_us": leg64["face_kernel_us"],
                     "flop_per_launch": leg64["flop_per_launch"],
                     "algorithmic_bytes_per_launch": leg64["bytes_per_launch"]},
    }
    if not args.no_mixed:
        progress("40962-vertex vesicle, FP32-mixed")
        legm, Em, gm = run_leg(args, w, c0, "mixed", d_xi, dev, torch)
        legm["speedup_vs_fp64"] = legm["face_evals_per_s"] / leg64["face_evals_per_s"]
        legm["face_kernel_speedup_vs_fp64"] = leg64["face_kernel_us"] / legm["face_kernel_us"]
        legm["kernel"] = kernel_name("mixed", legm["launch"]["point_split"], legm["launch"]["flags"])
        legm["traffic"] = measured_traffic(legm["kernel"])
        out["mixed"] = legm
    if not args.no_ensemble:
        progress("batched ensemble")
        out["ensemble"] = run_ensemble_leg(args, dev, torch)
    if not args.no_serving:
        progress("serving stream")
        out["serving"] = run_serving_leg(args, dev, torch)
    if not args.no_c1:
        progress("642-vertex vesicle, Brownian steps")
        out["c1"] = run_c1_leg(args, dev, torch)
    if not args.no_sheet:
        progress("128^2 sheet, Langevin steps")
        out["sheet"] = run_c2_leg(args, dev, torch)
    if not args.no_effm:
        progress("effective-mass integrators")
        out["effm"] = run_effm_leg(args, dev, torch)
    if not args.no_points:
        progress("surface points")
        out["points"] = run_points_leg(args, dev, torch)
    if not args.no_cpu_baseline:
        progress("CPU baseline of the 40962-vertex vesicle")
        base, E_o, g_o = cpu_baseline(w, c0, args.cpu_budget)
        out["cpu_baseline"] = base

        def rms(a):
            return float(np.sqrt(np.mean(a ** 2)))
        out["fp64"]["vs_oracle"] = {"energy_rel": abs(E64 - E_o) / abs(E_o),
                                    "grad_rms_rel": rms(g64[:-3] - g_o[:-3]) / rms(g_o[:-3])}
        if "mixed" in out:
            out["mixed"]["vs_oracle"] = {"energy_rel": abs(Em - E_o) / abs(E_o),
                                         "grad_rms_rel": rms(gm[:-3] - g_o[:-3]) / rms(g_o[:-3])}
    progress("done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
