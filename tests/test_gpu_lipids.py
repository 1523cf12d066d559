"""GPU parity of lipid compositions: the per-face, per-type lipid areas the
library packs at setg0 are bitwise the oracle's (both restate glibc rand()
after srand(seed) and the reference's packing loops, src/interpLib.C:3127-3196
and 3286-3358), and the energy and gradient with a three-type composition
agree with the oracle to 1e-12 (FP64) and 1e-6 / 1e-5 (mixed) on a vesicle,
a vesicle with valence defects and a periodic sheet.  A composition read
from a file of "lipid" directives equals the same composition set from arrays.
"""
import sys

import numpy as np
import pytest

try:  # load torch's HIP runtime before libumcp_core.so so both share one
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from oracle import oracle as O
from umcp_amd import surface as S
from umcp_amd import workloads

pytestmark = pytest.mark.gpu

# POPC, DOPE, CHOL of the reference's fixed library (c0 1/A, area per lipid A^2)
MIX = dict(c0=[-1.0 / 87.3, -1.0 / 29.4, -1.0 / 27.2], apl=[68.3, 63.4, 31.0],
           inner=[3.0, 2.0, 1.0], outer=[5.0, 0.0, 2.0])
SEED = 11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(w, precision=S.PREC_FP64, comp=MIX, seed=SEED):
    g = S.Surface(w.mesh_path)
    g.set_params(**{**w.params, "precision": precision})
    g.generatePlan()
    o = O.OracleSurface(w.mesh_path)
    o.set_params(**{k: v for k, v in w.params.items() if k != "precision"})
    g.set_lipids(comp["c0"], comp["apl"], comp["inner"], comp["outer"], seed)
    o.set_lipids(comp["c0"], comp["apl"], comp["inner"], comp["outer"], seed)
    g.setg0(w.r0)
    o.setg0(w.r0)
    return g, o


def check(g, o, r, e_tol, g_tol, tag):
    E_o, g_o = o.energy(r), o.grad(r)
    E_g, g_g = g.energy_grad(r)
    rel_e = abs(E_g - E_o) / abs(E_o)
    rel_g = np.linalg.norm(g_g[:-3] - g_o[:-3]) / np.linalg.norm(g_o[:-3])
    say("[lipids]", tag, "E", E_g, E_o, "rel", rel_e, "grad rel", rel_g)
    assert rel_e <= e_tol and rel_g <= g_tol


WORKLOADS = {
    "vesicle_642": workloads.vesicle_642,
    "defect_vesicle": lambda: workloads.defect_vesicle(),
    "defect_sheet": lambda: workloads.defect_sheet(),
}


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_lipid_types_and_energy(name):
    w = WORKLOADS[name]()
    say("[lipids] packing", name)
    g, o = make_pair(w)
    for f in range(g.nfaces):
        lo_g, li_g = g.face_lipid_types(f)
        lo_o, li_o = o.tri_lipid_types(o.face_info(f)["tri"])
        assert np.array_equal(lo_g, lo_o) and np.array_equal(li_g, li_o), f
        assert g.face_lipids(f) == o.tri_lipids(o.face_info(f)["tri"])
    check(g, o, w.r, 1e-12, 1e-12, name + " fp64")


@pytest.mark.parametrize("name", ["vesicle_642", "defect_vesicle"])
def test_lipid_energy_mixed(name):
    w = WORKLOADS[name]()
    g, o = make_pair(w, precision=S.PREC_MIXED)
    check(g, o, w.r, 1e-6, 1e-5, name + " mixed")


def test_composition_file_equals_arrays(tmp_path):
    w = workloads.vesicle_642()
    path = tmp_path / "lipids.inp"
    path.write_text("lipid library LIPX 55.5 -0.02\n"
                    "lipid inner POPC 3\nlipid inner DOPE 2\nlipid outer CHOL 2\n"
                    "lipid outer POPC 5\nlipid inner LIPX 1\n")
    a = S.Surface(w.mesh_path)
    a.set_params(**w.params)
    a.generatePlan()
    a.readLipidComposition(str(path), SEED)
    a.setg0(w.r0)
    # types in order of first mention: POPC, DOPE, CHOL, LIPX
    comp = dict(c0=[-1.0 / 87.3, -1.0 / 29.4, -1.0 / 27.2, -0.02], apl=[68.3, 63.4, 31.0, 55.5],
                inner=[3.0, 2.0, 0.0, 1.0], outer=[5.0, 0.0, 2.0, 0.0])
    b, o = make_pair(w, comp=comp)
    for f in range(a.nfaces):
        assert all(np.array_equal(x, y) for x, y in zip(a.face_lipid_types(f), b.face_lipid_types(f)))
    Ea, ga = a.energy_grad(w.r)
    Eb, gb = b.energy_grad(w.r)
    assert Ea == Eb and np.array_equal(ga, gb)
    check(a, o, w.r, 1e-12, 1e-12, "file composition")


def test_bad_compositions_are_refused(tmp_path):
    w = workloads.vesicle_642()
    g = S.Surface(w.mesh_path)
    g.set_params(**w.params)
    g.generatePlan()
    path = tmp_path / "bad.inp"
    path.write_text("lipid library BAD 0.0 -0.02\nlipid inner BAD 1\nlipid outer POPC 1\n")
    with pytest.raises(S.UmcpError):
        g.readLipidComposition(str(path), SEED)
    path.write_text("lipid inner NOSUCH 1\n")
    with pytest.raises(S.UmcpError):
        g.readLipidComposition(str(path), SEED)
    with pytest.raises(S.UmcpError):
        g.set_lipids([0.0], [-1.0], [1.0], [1.0], SEED)
    # an area per lipid so small that a face would hold more than 1e7 lipids
    g.set_lipids([0.0, 0.01], [68.3, 1e-9], [1.0, 1.0], [1.0, 1.0], SEED)
    with pytest.raises(S.UmcpError):
        g.setg0(w.r0)
    # a sane composition afterwards works
    g.set_lipids(MIX["c0"], MIX["apl"], MIX["inner"], MIX["outer"], SEED)
    g.setg0(w.r0)
    E, _ = g.energy_grad(w.r)
    assert np.isfinite(E)
