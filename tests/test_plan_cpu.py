"""The library's host half (mesh parsing, orientation, face plan, shared
tables) against the oracle, on the CPU: no GPU call is made here.

The reference orders each vertex ring from the triangle senses alone
(src/interpLib.C:1781-2020, 7515-7600), so rings written in any order must
give the same plan; both implementations are run on rings shuffled at random.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from umcp_amd import meshgen


def _host_plan(L, path):
    h = C.c_void_p()
    rc = L.umcp_load_mesh(path.encode(), C.byref(h))
    assert rc == 0, L.umcp_last_error(h)
    assert L.umcp_prepare(h) == 0, L.umcp_last_error(h)
    return h


def _points(L, h, nf):
    out = []
    for f in range(nf):
        ncp = C.c_int()
        cp = (C.c_int * 32)()
        assert L.umcp_face_points(h, f, C.byref(ncp), cp) == 0
        out.append(list(cp)[:ncp.value])
    return out


def _tables(L, h):
    res = []
    for c in range(L.umcp_class_count(h)):
        val, ncp, npt, nf = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.umcp_class_info(h, c, C.byref(val), C.byref(ncp), C.byref(npt), C.byref(nf))
        tab = np.zeros(npt.value * 6 * ncp.value)
        w = np.zeros(npt.value)
        L.umcp_class_table(h, c, tab.ctypes.data_as(C.POINTER(C.c_double)),
                           w.ctypes.data_as(C.POINTER(C.c_double)))
        res.append((val.value, tab.reshape(npt.value, 6, ncp.value), w))
    return res


MESHES = {
    "ico2": lambda: meshgen.icosphere(2, 50.0),
    "ico3": lambda: meshgen.icosphere(3, 100.0),
    "sheet12": lambda: meshgen.hex_sheet(12, 12, 100.0),
    "defect57": lambda: meshgen.defect_sphere(2, 100.0, 4, 7),   # valence 5 and 7
    "defect_sheet": lambda: meshgen.defect_sheet(16, 400.0, 3, 9),   # ... across periodic images
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("shuffle", [False, True])
def test_plan_matches_oracle(core_lib, mesh_dir, name, shuffle):
    mesh = MESHES[name]()
    if shuffle:
        rng = np.random.default_rng(len(name))
        mesh.rings = [list(rng.permutation(r)) for r in mesh.rings]
    path = str(mesh_dir / ("%s_%d.mesh" % (name, shuffle)))
    meshgen.write_mesh(mesh, path)
    L = core_lib
    h = _host_plan(L, path)
    try:
        o = O.OracleSurface(path)
        nr, ni = C.c_int(), C.c_int()
        nf = L.umcp_nfaces(h, C.byref(nr), C.byref(ni))
        assert (nf, nr.value, ni.value) == (o.nfaces, o.nreg, o.nirr)
        pts = _points(L, h, nf)
        for f in range(nf):
            assert pts[f] == o.face_info(f)["cp"], f
        for val, tab, w in _tables(L, h):
            f = next(f for f in range(nf) if o.face_info(f)["ncp"] == val + 6)
            for p in range(tab.shape[0]):
                w6, wt, _, _ = o.face_table(f, p)
                scale = max(1.0, np.abs(w6).max())
                assert np.abs(w6 - tab[p]).max() < 1e-13 * scale
                assert abs(wt - w[p]) < 1e-15
    finally:
        L.umcp_destroy(h)


def test_orientation_rule_ico(core_lib, mesh_dir):
    """Rings end up oriented so that (v, ring[e+1], ring[e]) is positively
    oriented relative to the lowest triangle (0, j, k) as the reference sets
    it; for the generated icosphere that is the inward sense, so the limit
    surface normal Ru x Rv points towards the centre."""
    mesh = meshgen.icosphere(2, 50.0)
    path = str(mesh_dir / "ico2_orient.mesh")
    meshgen.write_mesh(mesh, path)
    o = O.OracleSurface(path)
    j = o.neighbors(0)
    # consecutive ring entries are neighbours of each other
    for v in range(o.nv):
        ring = o.neighbors(v)
        for e in range(len(ring)):
            assert ring[(e + 1) % len(ring)] in o.neighbors(ring[e])
    assert j[0] == mesh.rings[0][0]


def test_illegal_meshes(core_lib, mesh_dir):
    L = core_lib
    # valence 8 is rejected at load (src/interpLib.C:1287-1291)
    bad = mesh_dir / "val8.mesh"
    bad.write_text("3D\n10 0 0\n0 10 0\n0 0 10\n0 0 0 0 8 1 2 3 4 5 6 7 8\n")
    h = C.c_void_p()
    assert L.umcp_load_mesh(str(bad).encode(), C.byref(h)) == -1
    assert b"valence" in L.umcp_last_error(h)
    L.umcp_destroy(h)
    # an irregular vertex next to a lower-index valence-6 owner is illegal
    # (src/interpLib.C:5500-5509): renumber an L1 icosphere so a regular vertex is 0
    m = meshgen.icosphere(1, 10.0)
    reg = next(v for v in range(m.nv) if len(m.rings[v]) == 6)
    perm = list(range(m.nv))
    perm[0], perm[reg] = perm[reg], perm[0]
    inv = np.argsort(perm)
    m2 = meshgen.Mesh(m.r[perm], [[int(inv[x]) for x in m.rings[perm[v]]] for v in range(m.nv)], m.pbc)
    path = str(mesh_dir / "illegal.mesh")
    meshgen.write_mesh(m2, path)
    h = C.c_void_p()
    assert L.umcp_load_mesh(path.encode(), C.byref(h)) == 0
    assert L.umcp_prepare(h) == -1
    assert b"illegal" in L.umcp_last_error(h)
    L.umcp_destroy(h)
    with pytest.raises(ValueError):
        O.OracleSurface(path)
    # a missing file
    h = C.c_void_p()
    assert L.umcp_load_mesh(b"/nonexistent/file.mesh", C.byref(h)) == -1
    L.umcp_destroy(h)
