"""libumcp_core.so loads and exports every entry point include/umcp.h declares
(CPU only; no compute call is made)."""
import ctypes as C
import os
import re

from umcp_amd import surface

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "umcp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(umcp_[a-z0-9_]+)\s*\(", txt)))


def test_exports_match_header(core_lib):
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(core_lib, n), n
    assert sorted(surface.EXPORTS) == names


def test_link_check(core_lib):
    assert core_lib.umcp_link_check() == 1


def test_null_handles(core_lib):
    L = core_lib
    assert L.umcp_nv(None) < 0
    assert L.umcp_plan(None) < 0
    assert L.umcp_energy(None, None, None) < 0
    assert L.umcp_last_error(None) == b"null handle"
    L.umcp_destroy(None)


def test_state_errors_without_plan(core_lib, mesh_dir):
    from umcp_amd import meshgen
    m = meshgen.icosphere(1, 10.0)
    path = str(mesh_dir / "ico1.mesh")
    meshgen.write_mesh(m, path)
    L = core_lib
    h = C.c_void_p()
    assert L.umcp_load_mesh(path.encode(), C.byref(h)) == 0
    r = (C.c_double * (3 * m.nv + 3))()
    E = C.c_double()
    assert L.umcp_energy(h, r, C.byref(E)) == -6   # UMCP_ERR_STATE: plan first
    assert b"plan" in L.umcp_last_error(h)
    L.umcp_destroy(h)
