"""Launch-shape autotuner for the face kernel.

The fused face kernel has four tunable choices: how many waves share the
Gauss points of a 64-face group (1..4), whether those waves also split the
control points (CPS form, splits 2..4), whether registers are bounded, and
(mixed precision) whether the basis contractions run on the matrix cores
(point splits 1, 2 and 4 there).  ``tune`` times every shape on the caller's
own surface with the library's HIP-event profiling and keeps the fastest; it
works on device buffers and never changes ``r`` (it evaluates energy +
gradient only).
"""
from __future__ import annotations

import itertools

SPLITS = (1, 2, 3, 4)
FLAGS = (0, 2)
CPS_SPLITS = (2, 3, 4)
CPS_FLAGS = (8,)
MFMA_SPLITS = (1, 2, 4)
MFMA_FLAGS = (4, 6)


def shapes(mixed: bool):
    out = list(itertools.product(SPLITS, FLAGS)) + list(itertools.product(CPS_SPLITS, CPS_FLAGS))
    if mixed:
        out += list(itertools.product(MFMA_SPLITS, MFMA_FLAGS))
    return out


def tune(surface, d_r, d_E, d_g, stream=None, reps: int = 10, sync=None, mixed: bool = False,
         rounds: int = 3):
    """Returns ((split, flags), {shape: microseconds}) and applies the best.

    The shapes lie within a few microseconds of each other, about the size of
    one short measurement's noise and of the clock ramp of an idle device, so
    the device is warmed first and every shape is timed in `rounds` rounds
    taken in turn; a shape's time is its fastest round's mean."""
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    cand = list(shapes(mixed))
    for split, flags in cand:                             # first launches: code load
        surface.set_launch(split, flags)
        surface.energy_grad_dev(d_r, d_E, d_g, stream)
    for _ in range(4 * reps):                             # clocks up
        surface.energy_grad_dev(d_r, d_E, d_g, stream)
    sync()
    times = {}
    for _ in range(rounds):
        for split, flags in cand:
            surface.set_launch(split, flags)
            surface.profile_begin()
            for _ in range(reps):
                surface.energy_grad_dev(d_r, d_E, d_g, stream)
            sync()
            ms, n = surface.profile_end()
            t = 1e3 * ms / max(n, 1)
            times[(split, flags)] = min(t, times.get((split, flags), t))
    best = min(times, key=times.get)
    surface.set_launch(*best)
    return best, times


# batched ensembles: waves per face group, register bound, persistent grid
# (0 = one workgroup per item) and item order
ENSEMBLE_SHAPES = (tuple((sp, fl, grid, order)
                         for sp in (2, 4) for fl in (0, 2) for grid in (0, 1024, 2048) for order in (0, 1))
                   + tuple((sp, 8, grid, order)
                           for sp in (2, 3, 4) for grid in (0, 1024, 2048) for order in (0, 1)))
# mixed precision adds the matrix-core path
ENSEMBLE_SHAPES_MIXED = ENSEMBLE_SHAPES + tuple((sp, fl, grid, order)
                                                for sp in (2, 4) for fl in (4, 6) for grid in (0, 1024)
                                                for order in (0, 1))


def tune_ensemble(ens, surfaces, d_rs, torch, stream, steps: int = 8, shapes=ENSEMBLE_SHAPES):
    """Times `steps` captured Brownian steps of the whole ensemble per launch
    shape and keeps the fastest.  Positions and step counters are restored
    afterwards, so tuning leaves the trajectory untouched.
    Returns (best shape, {shape: ms per step})."""
    saved = [d.clone() for d in d_rs]
    torch.cuda.synchronize()
    times = {}
    sptr = stream.cuda_stream
    for shape in shapes:
        ens.set_launch(*shape)
        ens.bd_steps(steps, steps, sptr)          # build, capture, first replay
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        ens.bd_steps(steps, steps, sptr)
        end.record(stream)
        torch.cuda.synchronize()
        times[shape] = start.elapsed_time(end) / steps
    best = min(times, key=times.get)
    ens.set_launch(*best)
    for d, s in zip(d_rs, saved):
        d.copy_(s)
    for surf in surfaces:
        surf.set_step(0)
    torch.cuda.synchronize()
    return best, times
