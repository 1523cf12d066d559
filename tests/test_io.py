"""On-disk formats (no GPU needed): the .mesh writer and restart files.

- .mesh (surface::saveSurface, src/write_structures.C:473-540): a saved mesh
  loads back with the same positions (bitwise, %.17g), valences and
  neighbour rings, and saving it again gives the same bytes;
- limit-surface frames (surface::writeLimitingSurface): three corner points
  per face; on a closed mesh every point is the limit position of one of the
  face's vertices (the oracle's Loop-mask limit points) and each vertex's
  point appears once per incident face;
- restart (src/restart.C:7-168): positions and momenta round-trip bitwise;
  the reference's own restart fixture (examples/test/energy_conservation/
  start.save: a 120-vertex planar mesh, positions only, one complex section,
  no seed line) reads with the values an independent parse of its rows gives.
"""
import ctypes as C
import os

import numpy as np
import pytest

from umcp_amd import surface as S
from umcp_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "energy_conservation_start.save")


def lib():
    return S.load_library()


def load(path):
    h = C.c_void_p()
    rc = lib().umcp_load_mesh(path.encode(), C.byref(h))
    assert rc == 0, lib().umcp_last_error(h)
    return h


def topology(h):
    L = lib()
    nv = L.umcp_nv(h)
    r = np.zeros(3 * nv + 3)
    L.umcp_get(h, S._dp(r))
    return nv, r


def rings(path):
    out = {}
    with open(path) as fp:
        for line in fp.readlines()[4:]:
            if line.startswith("ntri"):
                break
            f = line.split()
            out[int(f[0])] = (int(f[4]), [int(x) for x in f[5:]])
    return out


@pytest.mark.parametrize("kind", ["vesicle", "sheet"])
def test_mesh_round_trip(tmp_path, kind):
    w = workloads.vesicle_642() if kind == "vesicle" else workloads.sheet(32, seed=3, name="io_sheet32")
    h = load(w.mesh_path)
    nv, r_file = topology(h)
    r = w.r.copy()
    p1, p2 = str(tmp_path / "a.mesh"), str(tmp_path / "b.mesh")
    assert lib().umcp_save_mesh(h, S._dp(r), p1.encode()) == 0
    h2 = load(p1)
    nv2, r2 = topology(h2)
    assert nv2 == nv
    assert np.array_equal(r2[:-3], r[:-3])
    ra, rb = rings(w.mesh_path), rings(p1)
    for v in range(nv):
        assert ra[v][0] == rb[v][0] and sorted(ra[v][1]) == sorted(rb[v][1])
    assert lib().umcp_save_mesh(h2, None, p2.encode()) == 0
    assert open(p1).read() == open(p2).read()
    lib().umcp_destroy(h)
    lib().umcp_destroy(h2)


def test_restart_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    nv = 50
    r = rng.standard_normal(3 * nv + 3) * 100
    r[-3:] = [1.0, 1.0, 1.25]
    pp = rng.standard_normal(3 * nv) * 1e-3
    p = str(tmp_path / "x.save")
    S.write_restart(p, r, pp, seed=1234)
    got = S.read_restart(p, nv)
    assert np.array_equal(got["r"], r) and np.array_equal(got["pp"], pp)
    assert np.array_equal(got["alpha"], r[-3:]) and got["seed"] == 1234
    S.write_restart(p, r)
    got = S.read_restart(p, nv)
    assert np.array_equal(got["r"], r) and got["pp"] is None


def test_reference_restart_fixture():
    lines = open(FIXTURE).read().splitlines()
    nv = 120
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:nv + 2]])
    got = S.read_restart(FIXTURE, nv)
    assert np.array_equal(got["alpha"], [1.0, 1.0, 1.0])
    assert np.array_equal(got["r"].reshape(-1, 3), rows)
    assert got["pp"] is None and got["seed"] == -1
    assert np.array_equal(got["r"][:3], [-215.706610, -44.350899, -5.509793])
    with pytest.raises(S.UmcpError):
        S.read_restart(FIXTURE, 200)   # more rows than the file holds


def test_limit_surface_frames(tmp_path):
    from oracle import oracle as O
    w = workloads.vesicle_642()
    h = load(w.mesh_path)
    assert lib().umcp_prepare(h) == 0
    p = str(tmp_path / "limit.xyz")
    rng = np.random.default_rng(9)
    r = w.r.copy()
    r[:-3] += 0.2 * rng.standard_normal(r.size - 3)
    assert lib().umcp_write_limit_surface(h, S._dp(r), p.encode(), 0) == 0
    assert lib().umcp_write_limit_surface(h, S._dp(r), p.encode(), 1) == 0
    lines = open(p).read().splitlines()
    nf = 1280
    assert lines[0] == str(3 * nf) and lines[1] == "limiting surface"
    assert len(lines) == 2 * (2 + 3 * nf) and lines[2 + 3 * nf] == str(3 * nf)
    pts = np.array([[float(x) for x in ln.split()[1:]] for ln in lines[2:2 + 3 * nf]])
    o = O.OracleSurface(w.mesh_path)
    L = o.limit_points(r).reshape(-1, 3)
    d = np.linalg.norm(pts[:, None, :] - L[None, :, :], axis=2)
    nearest = d.argmin(axis=1)
    assert d.min(axis=1).max() < 1e-5           # %lf keeps six decimals
    counts = np.bincount(nearest, minlength=o.nv)
    val = np.array([len(o.neighbors(v)) for v in range(o.nv)])
    assert np.array_equal(counts, val)
    lib().umcp_destroy(h)


def test_limit_surface_psf(tmp_path):
    """The PSF topology of the limit frames (writeLimitingSurfacePSF at plim = 1,
    src/write_structures.C:544-655; writePSF, src/mutil.C:601-676): the
    column layout, one atom per frame point, and bonds that join the three
    corners of each face in the frame written next to it."""
    w = workloads.vesicle_642()
    h = load(w.mesh_path)
    assert lib().umcp_prepare(h) == 0
    p = str(tmp_path / "limit.psf")
    x = str(tmp_path / "limit.xyz")
    assert lib().umcp_write_limit_psf(h, p.encode()) == 0
    r = w.r.copy()
    assert lib().umcp_write_limit_surface(h, S._dp(r), x.encode(), 0) == 0
    lines = open(p).read().split("\n")
    nf = 1280
    n = 3 * nf
    assert lines[0] == "PSF EXT" and lines[1] == ""
    assert lines[2] == "    3840 !NATOM"
    tail = "C        C        C        0.000000        1.0000           0   0.00000      0.000000E-01"
    assert lines[3] == "         1 SEGN     1        " + tail
    assert lines[3 + 1233] == "      1234 SEGN     1234     " + tail
    for i in range(n):
        f = lines[3 + i].split()
        assert f[0] == f[2] == str(i + 1) and f[1] == "SEGN"
    assert lines[3 + n] == "    3840 !NBOND: nbonds"
    body = lines[4 + n:]
    assert all(len(ln) == 80 for ln in body[:-1]) and body[-1] == ""   # 4 bonds of 2 x 10 columns per line
    tok = np.array("".join(body).split(), dtype=int).reshape(-1, 2)
    assert len(tok) == 3 * nf
    t = np.arange(nf)
    want = np.stack([np.stack([3 * t + 1, 3 * t + 3], 1), np.stack([3 * t + 1, 3 * t + 2], 1),
                     np.stack([3 * t + 2, 3 * t + 3], 1)], 1).reshape(-1, 2)
    assert np.array_equal(tok, want)
    # each bond joins two distinct corners of one face of the frame: an edge
    pts = np.array([[float(v) for v in ln.split()[1:]] for ln in open(x).read().splitlines()[2:2 + n]])
    d = np.linalg.norm(pts[tok[:, 0] - 1] - pts[tok[:, 1] - 1], axis=1)
    assert d.min() > 0 and d.max() < 3 * np.median(d)
    lib().umcp_destroy(h)
