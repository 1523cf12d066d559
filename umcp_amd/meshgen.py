"""Seeded synthetic meshes in the reference ``.mesh`` format.

Mirrors the reference's fixture tools:

* ``icosphere``: the 12-vertex icosahedron of ``src/icosahedron.C:8-79`` (same
  vertex layout and neighbour rings, box ``3R``), refined by Loop subdivision
  that keeps the old vertices at indices ``0..nvo-1`` and numbers the new edge
  vertices in the order ``src/interpLib.C:2607-2700`` visits edges (vertex
  ascending, then ring order, for ``j > i``), with the Loop / Cirak vertex and
  edge rules, then radially projected to radius ``R``.
* ``hex_sheet``: an exactly periodic triangulated hex lattice, every vertex of
  valence 6, with a height field that is a sum of Fourier modes of amplitude
  proportional to ``q^-2``.

The writer follows ``surface::saveSurface`` (``src/write_structures.C:
473-540``) but prints coordinates with 17 significant digits instead of
``%lf`` so a mesh round-trips exactly.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    """Vertex positions (nv x 3), ordered neighbour rings, periodic box rows."""

    r: np.ndarray
    rings: list
    pbc: np.ndarray = field(default_factory=lambda: np.diag([1e10, 1e10, 1e10]))

    @property
    def nv(self) -> int:
        return int(self.r.shape[0])

    def valence(self) -> np.ndarray:
        return np.array([len(x) for x in self.rings], dtype=np.int32)

    def triangles(self) -> np.ndarray:
        """Oriented triangles (i, ring[e], ring[e+1]) owned by their lowest vertex."""
        out = []
        for i, ring in enumerate(self.rings):
            n = len(ring)
            for e in range(n):
                j, k = ring[e], ring[(e + 1) % n]
                if i < j and i < k:
                    out.append((i, j, k))
        return np.array(out, dtype=np.int64)


def write_mesh(mesh: Mesh, path: str, with_triangles: bool = False) -> None:
    with open(path, "w") as fh:
        fh.write("3D synthetic surface\n")
        for row in mesh.pbc:
            fh.write("%.17g %.17g %.17g\n" % tuple(row))
        for v in range(mesh.nv):
            ring = mesh.rings[v]
            fh.write("%d %.17g %.17g %.17g %d %s\n" % (
                v, mesh.r[v, 0], mesh.r[v, 1], mesh.r[v, 2], len(ring),
                " ".join(str(int(x)) for x in ring)))
        if with_triangles:
            tris = mesh.triangles()
            fh.write("ntri %d\n" % len(tris))
            for i, j, k in tris:
                a, b = (j, k) if j < k else (k, j)
                fh.write("%d %d %d\n" % (i, a, b))


def read_mesh(path: str) -> Mesh:
    with open(path) as fh:
        lines = fh.read().splitlines()
    vec3 = lines[0][:2].lower() == "3d"
    pbc = np.zeros((3, 3))
    pbc[0] = [float(x) for x in lines[1].split()[:3]]
    pbc[1] = [float(x) for x in lines[2].split()[:3]]
    pos = 3
    if vec3:
        pbc[2] = [float(x) for x in lines[3].split()[:3]]
        pos = 4
    else:
        pbc[2] = [0, 0, 1e6]
    if pos < len(lines) and lines[pos].startswith("sense"):
        pos += 1
    r, rings = [], []
    for ln in lines[pos:]:
        if ln.startswith("ntri"):
            break
        f = ln.split()
        if len(f) < 5:
            continue
        val = int(f[4])
        r.append([float(f[1]), float(f[2]), float(f[3])])
        rings.append([int(x) for x in f[5:5 + val]])
    return Mesh(np.array(r, dtype=np.float64), rings, pbc)


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------

_ICO_RINGS = [
    [2, 3, 4, 5, 6], [7, 8, 9, 10, 11],
    [0, 6, 3, 7, 11], [0, 2, 4, 8, 7], [0, 3, 5, 9, 8], [0, 4, 6, 10, 9], [0, 5, 2, 11, 10],
    [1, 8, 11, 2, 3], [1, 9, 7, 3, 4], [1, 10, 8, 4, 5], [1, 11, 9, 5, 6], [1, 7, 10, 6, 2],
]


def icosahedron(R: float = 1.0) -> Mesh:
    r = np.zeros((12, 3))
    r[0] = [0, 0, R]
    r[1] = [0, 0, -R]
    phi1 = math.atan(0.5)
    theta = 2 * math.pi / 5.0
    phase = 2 * math.pi / 10.0
    for i in range(5):
        r[2 + i] = [R * math.sin(i * theta) * math.cos(phi1),
                    R * math.cos(i * theta) * math.cos(phi1), R * math.sin(phi1)]
        r[7 + i] = [R * math.sin(i * theta + phase) * math.cos(phi1),
                    R * math.cos(i * theta + phase) * math.cos(phi1), -R * math.sin(phi1)]
    rings = [_ccw_ring(r, v, list(ring)) for v, ring in enumerate(_ICO_RINGS)]
    return Mesh(r, rings, np.diag([3.0 * R] * 3))


def _ccw_ring(r: np.ndarray, v: int, ring: list) -> list:
    """Order a closed ring counter-clockwise about the outward radial direction."""
    n = r[v] / np.linalg.norm(r[v])
    t1 = np.cross(n, [0.3, 0.5, 0.81])
    if np.linalg.norm(t1) < 1e-8:
        t1 = np.cross(n, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    ang = [math.atan2(np.dot(r[j] - r[v], t2), np.dot(r[j] - r[v], t1)) for j in ring]
    order = [ring[i] for i in np.argsort(ang)]
    s = order.index(ring[0])
    return order[s:] + order[:s]


def loop_subdivide(mesh: Mesh, periodic: bool = False) -> Mesh:
    """One Loop step: old vertices keep their ids; edge vertices appended.
    periodic: the mesh tiles its (rectangular) box, and stencils take each
    neighbour at its nearest image."""
    nvo = mesh.nv
    rings = mesh.rings
    edge_id = {}
    nv = nvo
    for i in range(nvo):
        for j in rings[i]:
            if j > i:
                edge_id[(i, j)] = nv
                nv += 1

    def mid(a, b):
        return edge_id[(a, b) if a < b else (b, a)]

    r = np.zeros((nv, 3))
    if periodic:
        # rectangular periodic box: neighbours at their image nearest the
        # vertex the stencil is centred on, results wrapped into the box
        L = np.diag(mesh.pbc)

        def near(a, b):
            d = mesh.r[b] - mesh.r[a]
            return mesh.r[a] + (d - L * np.round(d / L))
    for v in range(nvo):
        val = len(rings[v])
        w = 3.0 / (8.0 * val)
        if periodic:
            r[v] = (1 - val * w) * mesh.r[v] + w * sum(near(v, j) for j in rings[v])
        else:
            r[v] = (1 - val * w) * mesh.r[v] + w * mesh.r[rings[v]].sum(axis=0)
    # triangle-fan successor: in ring order, (i, ring[e], ring[e+1]) is oriented
    for (a, b), m in edge_id.items():
        ra = rings[a]
        e = ra.index(b)
        c = ra[(e + 1) % len(ra)]
        d = ra[(e - 1) % len(ra)]
        if periodic:
            r[m] = 3.0 / 8.0 * (mesh.r[a] + near(a, b)) + 1.0 / 8.0 * (near(a, c) + near(a, d))
        else:
            r[m] = 3.0 / 8.0 * (mesh.r[a] + mesh.r[b]) + 1.0 / 8.0 * (mesh.r[c] + mesh.r[d])
    if periodic:
        r = r - L * np.floor(r / L)
    new_rings = [None] * nv
    for v in range(nvo):
        new_rings[v] = [mid(v, j) for j in rings[v]]
    for (a, b), m in edge_id.items():
        ra = rings[a]
        e = ra.index(b)
        c = ra[(e + 1) % len(ra)]   # triangle (a, b, c) oriented
        d = ra[(e - 1) % len(ra)]   # triangle (a, d, b) oriented
        # ring around m consistent with (a, b, c) orientation
        new_rings[m] = [a, mid(a, d), mid(b, d), b, mid(b, c), mid(a, c)]
    return Mesh(r, new_rings, mesh.pbc.copy())


def _orient_like_faces(mesh: Mesh) -> Mesh:
    """Reverse every ring if the first triangle's normal points inward."""
    i, j, k = mesh.triangles()[0]
    n = np.cross(mesh.r[j] - mesh.r[i], mesh.r[k] - mesh.r[i])
    if np.dot(n, mesh.r[i] + mesh.r[j] + mesh.r[k]) < 0:
        rings = [[ring[0]] + ring[1:][::-1] for ring in mesh.rings]
        return Mesh(mesh.r, rings, mesh.pbc)
    return mesh


def icosphere(level: int, R: float) -> Mesh:
    """Icosahedron refined ``level`` times by Loop subdivision, projected to R."""
    m = icosahedron(1.0)
    for _ in range(level):
        m = loop_subdivide(m)
    r = m.r / np.linalg.norm(m.r, axis=1, keepdims=True) * R
    return _orient_like_faces(Mesh(r, m.rings, np.diag([3.0 * R] * 3)))


def flip_edge(mesh: Mesh, a: int, b: int) -> Mesh:
    """Replaces edge a-b by the edge between its two opposite vertices.

    With triangles (a, b, c) and (a, d, b) on the edge, the new triangles are
    (a, d, c) and (b, c, d): a and b lose a neighbour, c and d gain one.
    """
    rings = [list(x) for x in mesh.rings]
    ra = rings[a]
    e = ra.index(b)
    c, d = ra[(e + 1) % len(ra)], ra[(e - 1) % len(ra)]
    ra.remove(b)
    rings[b].remove(a)
    rc = rings[c]
    rc.insert(rc.index(a) + 1, d)        # c: ... a, d, b ...
    rd = rings[d]
    rd.insert(rd.index(b) + 1, c)        # d: ... b, c, a ...
    return Mesh(mesh.r.copy(), rings, mesh.pbc.copy())


def defect_sphere(level: int, R: float, nflip: int, seed: int) -> Mesh:
    """A sphere with valence-5 and valence-7 vertices: nflip edges between
    valence-6 vertices of icosphere(level) are flipped (each makes two
    valence-5 and two valence-7 vertices), chosen at random among edges whose
    four vertices are untouched and of valence 6; one Loop step then
    separates the irregular vertices, and the result is projected to R."""
    m = icosphere(level, 1.0)
    rng = np.random.default_rng(seed)
    touched = set()
    edges = [(i, j) for i in range(m.nv) for j in m.rings[i] if j > i]
    order = rng.permutation(len(edges))
    done = 0
    for idx in order:
        if done == nflip:
            break
        a, b = edges[idx]
        ra = m.rings[a]
        e = ra.index(b)
        c, d = ra[(e + 1) % len(ra)], ra[(e - 1) % len(ra)]
        quad = (a, b, c, d)
        near = set(quad)
        for v in quad:
            near.update(m.rings[v])
        if any(len(m.rings[v]) != 6 for v in quad) or near & touched:
            continue
        m = flip_edge(m, a, b)
        touched.update(near)
        done += 1
    if done < nflip:
        raise ValueError("could not place %d separated flips" % nflip)
    m = loop_subdivide(m)
    r = m.r / np.linalg.norm(m.r, axis=1, keepdims=True) * R
    return _orient_like_faces(Mesh(r, m.rings, np.diag([3.0 * R] * 3)))


def defect_sheet(n: int, Lx: float, nflip: int, seed: int) -> Mesh:
    """A periodic n x n hex sheet with valence-5 and valence-7 vertices: nflip
    separated edge flips, then one periodic Loop step (2n x 2n vertices)."""
    m = hex_sheet(n, n, Lx)
    rng = np.random.default_rng(seed)
    touched = set()
    edges = [(i, j) for i in range(m.nv) for j in m.rings[i] if j > i]
    done = 0
    for idx in rng.permutation(len(edges)):
        if done == nflip:
            break
        a, b = edges[idx]
        ra = m.rings[a]
        e = ra.index(b)
        c, d = ra[(e + 1) % len(ra)], ra[(e - 1) % len(ra)]
        near = {a, b, c, d}
        for v in (a, b, c, d):
            near.update(m.rings[v])
        if near & touched:
            continue
        m = flip_edge(m, a, b)
        touched.update(near)
        done += 1
    if done < nflip:
        raise ValueError("could not place %d separated flips" % nflip)
    m = loop_subdivide(m, periodic=True)
    # move the first valence-7 vertex next to the box corner, so that its
    # faces reach across both periodic boundaries
    L = np.diag(m.pbc)
    v7 = int(np.flatnonzero(m.valence() == 7)[0])
    shift = 0.15 * (Lx / (2 * n)) - m.r[v7]
    shift[2] = 0.0
    r = m.r + shift
    r = r - L * np.floor(r / L)
    return Mesh(r, m.rings, m.pbc.copy())


# --------------------------------------------------------------------------
# periodic hex sheet
# --------------------------------------------------------------------------

def hex_sheet(nx: int, ny: int, Lx: float, Lz: float = 1000.0) -> Mesh:
    """nx x ny triangular lattice (ny even) in a periodic Lx x (ny a sqrt3/2) box."""
    if ny % 2:
        raise ValueError("ny must be even for a periodic hex lattice")
    a = Lx / nx
    Ly = ny * a * math.sqrt(3.0) / 2.0
    r = np.zeros((nx * ny, 3))
    rings = []
    for iy in range(ny):
        for ix in range(nx):
            v = iy * nx + ix
            r[v] = [(ix + 0.5 * (iy % 2)) * a, iy * a * math.sqrt(3.0) / 2.0, 0.0]
            o = iy % 2
            nb = [(ix + 1, iy), (ix + o, iy + 1), (ix - 1 + o, iy + 1),
                  (ix - 1, iy), (ix - 1 + o, iy - 1), (ix + o, iy - 1)]
            rings.append([(y % ny) * nx + (x % nx) for x, y in nb])
    return Mesh(r, rings, np.diag([Lx, Ly, Lz]))


def fourier_heights(mesh: Mesh, kc_over_kT: float, nmax: int, seed: int) -> np.ndarray:
    """Height field sum_q A_q cos(q.r + phi_q), A_q = 1/(q^2 sqrt(kc/kT Lx Ly))."""
    rng = np.random.default_rng(seed)
    Lx, Ly = mesh.pbc[0, 0], mesh.pbc[1, 1]
    h = np.zeros(mesh.nv)
    x, y = mesh.r[:, 0], mesh.r[:, 1]
    for nx in range(-nmax, nmax + 1):
        for ny in range(0, nmax + 1):
            if ny == 0 and nx <= 0:
                continue
            qx, qy = 2 * math.pi * nx / Lx, 2 * math.pi * ny / Ly
            q2 = qx * qx + qy * qy
            amp = 1.0 / (q2 * math.sqrt(kc_over_kT * Lx * Ly))
            h += amp * np.cos(qx * x + qy * y + rng.uniform(0, 2 * math.pi))
    return h


def flat_coords(mesh: Mesh) -> np.ndarray:
    """The reference ``r`` layout: 3 nv coordinates followed by the alpha triple."""
    return np.concatenate([mesh.r.reshape(-1), [1.0, 1.0, 1.0]])
