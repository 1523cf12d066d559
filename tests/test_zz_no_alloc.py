"""umcp_energy_grad_dev_noalloc: the no-allocation variant of
umcp_energy_grad_dev (its work buffers in the library's static scratch array)
against umcp_energy_grad_dev itself, through the C ABI.

Bar: bitwise equal energy and gradient (same kernels, same order).  Two calls
in a row, the second on a smaller surface, so the second reuses the array the
first filled; FP64 and mixed precision; restraints on and off; a surface too
large for the array is rejected with UMCP_ERR_ARG and a message, and leaves
the handle usable.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from umcp_amd import surface as S  # noqa: E402
from umcp_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _surface(w, precision, c0=None, restraints=None):
    g = S.Surface(w.mesh_path)
    g.set_params(precision=precision, **w.params)
    g.generatePlan()
    if c0 is not None:
        g.set_face_c0(c0)
    g.setg0(w.r0)
    if restraints:
        g.set_restraints(**restraints)
    return g


def _both(g, r):
    """(E, g) of the allocating entry point, then of the no-allocation one."""
    d_r = torch.tensor(r, device="cuda", dtype=torch.float64)
    out = []
    for fn in (g.energy_grad_dev, g.energy_grad_dev_noalloc):
        d_E = torch.full((1,), np.nan, device="cuda", dtype=torch.float64)
        d_g = torch.full((r.size,), np.nan, device="cuda", dtype=torch.float64)
        fn(d_r, d_E, d_g)
        torch.cuda.synchronize()
        out.append((float(d_E.cpu()[0]), d_g.cpu().numpy()))
    return out


@pytest.mark.parametrize("precision", [S.PREC_FP64, S.PREC_MIXED])
def test_noalloc_equals_allocating_entry_point(mesh_dir, precision):
    big = workloads.vesicle(4, 300.0, 1.5, seed=21, name="vesicle_2562_terms")
    small = workloads.vesicle_642()
    gb = _surface(big, precision, c0=None, restraints={"KA_global": 0.1, "KV": 0.004})
    gs = _surface(small, precision)
    for tag, g, w in (("2562 restraints", gb, big), ("642", gs, small), ("2562 again", gb, big)):
        say("[noalloc]", tag, "precision", precision)
        (E1, g1), (E2, g2) = _both(g, w.r)
        assert np.isfinite(E1) and np.isfinite(g1).all()
        assert E1 == E2, (tag, E1, E2)
        assert np.array_equal(g1, g2), (tag, np.abs(g1 - g2).max())


def test_noalloc_rejects_a_surface_beyond_the_array(mesh_dir):
    """The 163842-vertex icosphere has about 480k gradient records (one per
    distinct vertex of each 64-face group): 11.5 MB in FP64, beyond the 8 MiB
    array: rejected; in mixed precision (4-byte records, 5.8 MB) it fits and
    matches the allocating call."""
    w = workloads.vesicle(7, 2000.0, 2.0, seed=5, name="vesicle_163842_noalloc")
    g = _surface(w, S.PREC_FP64)
    d_r = torch.tensor(w.r, device="cuda", dtype=torch.float64)
    d_E = torch.zeros(1, device="cuda", dtype=torch.float64)
    d_g = torch.zeros(w.r.size, device="cuda", dtype=torch.float64)
    say("[noalloc] 163842-vertex icosphere FP64 (rejected)")
    with pytest.raises(S.UmcpError, match="scratch"):
        g.energy_grad_dev_noalloc(d_r, d_E, d_g)
    g.energy_grad_dev(d_r, d_E, d_g)
    torch.cuda.synchronize()
    assert np.isfinite(float(d_E.cpu()[0]))
    say("[noalloc] 163842-vertex icosphere mixed")
    gm = _surface(w, S.PREC_MIXED)
    (E1, g1), (E2, g2) = _both(gm, w.r)
    assert E1 == E2 and np.array_equal(g1, g2)
