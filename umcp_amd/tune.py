"""Launch-shape autotuner for the face kernel.

The fused face kernel has two tunable choices: how many waves share the
Gauss points of a 64-face group (1..4) and whether registers are bounded for
two waves per SIMD.  ``tune`` times every
shape on the caller's own surface with the library's HIP-event profiling and
keeps the fastest; it works on device buffers and never changes ``r`` (it
evaluates energy + gradient only).
"""
from __future__ import annotations

import itertools

SPLITS = (1, 2, 3, 4)
FLAGS = (0, 2)


def tune(surface, d_r, d_E, d_g, stream=None, reps: int = 10, sync=None):
    """Returns ((split, flags), {shape: microseconds}) and applies the best."""
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    times = {}
    for split, flags in itertools.product(SPLITS, FLAGS):
        surface.set_launch(split, flags)
        surface.energy_grad_dev(d_r, d_E, d_g, stream)   # first launch: code load
        sync()
        surface.profile_begin()
        for _ in range(reps):
            surface.energy_grad_dev(d_r, d_E, d_g, stream)
        sync()
        ms, n = surface.profile_end()
        times[(split, flags)] = 1e3 * ms / max(n, 1)
    best = min(times, key=times.get)
    surface.set_launch(*best)
    return best, times
