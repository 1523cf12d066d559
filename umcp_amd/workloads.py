"""The seeded synthetic workloads of the benchmark and parity tests.

Each builder writes its ``.mesh`` file (reference format) into a cache
directory and returns a ``Workload`` with the reference state positions
``r0`` (passed to setg0), the evaluation positions ``r``, and the physical
parameters in reference units (kcal/mol, Angstrom).

* ``vesicle_642``    icosahedron Loop-subdivided 3 times (642 vertices, 1280
  faces, 60 of them irregular), radius 100 A, isotropic N(0, 0.5^2) noise,
  kc = 20 kT, c0 = 0, KA = 0.215 kT/A^2 (T = 298 K).
* ``sheet_128``      periodic 128 x 128 hex lattice in a 1000 x 866 A box,
  heights = sum of Fourier modes with amplitude ~ q^-2 (kc = 20 kT, T = 310 K).
* ``vesicle_40962``  icosahedron subdivided 6 times, radius 1000 A, radial
  N(0, 2^2) noise, kc = 20 kT, per-face c0 ~ U[-0.01, 0.01] 1/A (Helfrich form),
  global area and volume restraints.

Reference state: the unperturbed mesh (r0) is the setg0 state; the energy is
evaluated at the perturbed r.
"""
from __future__ import annotations

import hashlib
import math
import os
import tempfile
from dataclasses import dataclass, field

import numpy as np

from . import meshgen, units

GENERATOR_VERSION = 1


@dataclass
class Workload:
    name: str
    mesh_path: str
    nv: int
    r0: np.ndarray
    r: np.ndarray
    params: dict
    face_c0: np.ndarray | None = None
    T: float = 298.0
    extra: dict = field(default_factory=dict)


def cache_dir() -> str:
    d = os.environ.get("UMCP_CACHE", os.path.join(tempfile.gettempdir(), "umcp_meshes"))
    os.makedirs(d, exist_ok=True)
    return d


def _write_once(mesh: meshgen.Mesh, name: str) -> str:
    path = os.path.join(cache_dir(), "%s_v%d.mesh" % (name, GENERATOR_VERSION))
    tmp = path + ".%d.tmp" % os.getpid()
    meshgen.write_mesh(mesh, tmp)
    os.replace(tmp, path)
    return path


def vesicle(level: int, R: float, sigma: float, seed: int, radial: bool = False,
            kc_kT: float = 20.0, KA_kT: float = 0.215, T: float = 298.0, name: str | None = None):
    mesh = meshgen.icosphere(level, R)
    path = _write_once(mesh, name or "ico_L%d_R%g" % (level, R))
    r0 = meshgen.flat_coords(mesh)
    rng = np.random.default_rng(seed)
    r = r0.copy()
    if radial:
        nrm = mesh.r / np.linalg.norm(mesh.r, axis=1, keepdims=True)
        r[:-3] += (nrm * rng.normal(0.0, sigma, (mesh.nv, 1))).reshape(-1)
    else:
        r[:-3] += rng.normal(0.0, sigma, 3 * mesh.nv)
    params = dict(kc=units.kc_from_kT(kc_kT, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0,
                  mode=0)
    return Workload(name or path, path, mesh.nv, r0, r, params, T=T)


def defect_vesicle(level: int = 2, R: float = 100.0, nflip: int = 4, sigma: float = 0.5, seed: int = 7,
                   kc_kT: float = 20.0, KA_kT: float = 0.215, T: float = 298.0) -> Workload:
    """A vesicle with valence-5 and valence-7 vertices (meshgen.defect_sphere:
    nflip edge flips on icosphere(level), one Loop step), isotropic noise."""
    mesh = meshgen.defect_sphere(level, R, nflip, seed)
    name = "defect_L%d_f%d_s%d" % (level, nflip, seed)
    path = _write_once(mesh, name)
    r0 = meshgen.flat_coords(mesh)
    rng = np.random.default_rng(seed + 1)
    r = r0.copy()
    r[:-3] += rng.normal(0.0, sigma, 3 * mesh.nv)
    params = dict(kc=units.kc_from_kT(kc_kT, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0,
                  mode=0)
    return Workload(name, path, mesh.nv, r0, r, params, T=T)


def defect_sheet(n: int = 16, Lx: float = 400.0, nflip: int = 3, seed: int = 9, kc_kT: float = 20.0,
                 T: float = 310.0, KA_kT: float = 0.215) -> Workload:
    """A periodic sheet with valence-5 and valence-7 vertices (meshgen.defect_sheet)
    under a Fourier height field: irregular faces with periodic images."""
    mesh = meshgen.defect_sheet(n, Lx, nflip, seed)
    name = "defect_sheet_%d_f%d_s%d" % (n, nflip, seed)
    path = _write_once(mesh, name)
    r0 = meshgen.flat_coords(mesh)
    h = meshgen.fourier_heights(mesh, kc_kT, 4, seed)
    r = r0.copy()
    r[2:-3:3] += h
    params = dict(kc=units.kc_from_kT(kc_kT, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0,
                  mode=0)
    return Workload(name, path, mesh.nv, r0, r, params, T=T)


def vesicle_642(seed: int = 1) -> Workload:
    return vesicle(3, 100.0, 0.5, seed, name="vesicle_642")


def vesicle_40962(seed: int = 3, with_c0: bool = True, with_restraints: bool = True) -> Workload:
    """The large vesicle; with_restraints: global area and volume restraints
    (``extra["restraints"]``: KA_global = 0.215 kT/A^2 on the total area, as
    the reference's GLOBAL_AREA build, which drops the per-point KA term, and
    KV = 0.0056 kcal/mol/A^3, the reference's water_KV default,
    src/interpLib.C:91), targets = the setg0 state."""
    w = vesicle(6, 1000.0, 2.0, seed, radial=True, name="vesicle_40962")
    if with_c0:
        w.params["mode"] = 1
        w.extra["c0_seed"] = seed + 100
    if with_restraints:
        w.extra["restraints"] = dict(KA_global=w.params["KA"], KV=0.0056)
        w.params["KA"] = 0.0
    return w


def face_c0_uniform(nfaces: int, seed: int, lo: float = -0.01, hi: float = 0.01) -> np.ndarray:
    return np.random.default_rng(seed).uniform(lo, hi, nfaces)


def sheet(n: int, Lx: float = 1000.0, seed: int = 2, kc_kT: float = 20.0, T: float = 310.0,
          KA_kT: float = 0.215, nmax: int = 6, name: str | None = None) -> Workload:
    mesh = meshgen.hex_sheet(n, n, Lx)
    path = _write_once(mesh, name or "sheet_%d" % n)
    r0 = meshgen.flat_coords(mesh)
    h = meshgen.fourier_heights(mesh, kc_kT, nmax, seed)
    r = r0.copy()
    r[2:-3:3] += h
    params = dict(kc=units.kc_from_kT(kc_kT, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0,
                  mode=0)
    return Workload(name or path, path, mesh.nv, r0, r, params, T=T)


def sheet_128(seed: int = 2) -> Workload:
    return sheet(128, seed=seed, name="sheet_128")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


@dataclass
class Member:
    """One surface of a batched ensemble: a template mesh scaled to radius R,
    with its own Brownian time step dt (s) and mean edge length h (A)."""
    level: int
    R: float
    kc_kT: float
    seed: int
    mesh_path: str
    r0: np.ndarray
    r: np.ndarray
    params: dict
    T: float = 298.0
    dt: float = 1e-9
    h: float = 0.0

    @property
    def nv(self) -> int:
        return (self.r.size - 3) // 3

    @property
    def nfaces(self) -> int:
        return 2 * self.nv - 4


def icosphere_template(level: int) -> tuple[str, np.ndarray, float]:
    """Mesh file, coordinates and mean edge length of a level-L icosphere of
    radius 100 A.  Closed meshes plan identically at any radius (no periodic
    images), so the members of an ensemble share one file per level and scale
    its positions."""
    mesh = meshgen.icosphere(level, 100.0)
    edges = [np.linalg.norm(mesh.r[v] - mesh.r[u]) for v in range(mesh.nv) for u in mesh.rings[v] if u > v]
    return _write_once(mesh, "ico_L%d_R100" % level), meshgen.flat_coords(mesh), float(np.mean(edges))


def ensemble(n: int = 256, seed: int = 4, levels=(2, 3, 4, 5), weights=(4, 3, 2, 1),
             R_range=(50.0, 500.0), kc_range=(10.0, 40.0), KA_kT: float = 0.215, T: float = 298.0,
             rel_sigma: float = 0.05, dt_max: float = 1e-9, h_ref: float = 13.0,
             dt_safety: float = 0.3) -> list[Member]:
    """The batched ensemble: n closed vesicles with seeded level (probabilities
    weights/sum: level 2 four times as likely as level 5), radius ~ U[R_range],
    kc ~ U[kc_range] kT, Helfrich energy form (mode B, c0 = 0) with the local
    area term KA = 0.215 kT/A^2, D = 1e-12 m^2/s.  (The reference's default
    lipid form gives every triangle at least one 68.3 A^2 lipid, which on the
    2-20 A edges of these meshes overweights bending by up to ~30x.)

    Choices of this workload (the specification leaves them open): the initial
    positional noise is N(0, (rel_sigma h)^2) with h the member's mean edge
    length, and each member's Brownian step is dt = min(dt_max, dt_safety
    dt_max (h / h_ref)^2).  The stiffest bending mode of a mesh relaxes at a
    rate ~ (D/kT) kc / h^2, so explicit Euler needs dt ~ h^2: with D = 1e-12
    m^2/s, dt = 1e-9 s is stable at the 13 A edges of the 642-vertex R = 100 A
    vesicle, while a level-5 mesh of R = 50 A (1.6 A edges) diverges even at
    2e-11 s."""
    rng = np.random.default_rng(seed)
    p = np.asarray(weights, dtype=float) / float(sum(weights))
    tmpl = {L: icosphere_template(L) for L in levels}
    out = []
    for i in range(n):
        L = int(rng.choice(levels, p=p))
        R = float(rng.uniform(*R_range))
        kc = float(rng.uniform(*kc_range))
        path, base, h100 = tmpl[L]
        h = h100 * R / 100.0
        r0 = base.copy()
        r0[:-3] *= R / 100.0
        r = r0.copy()
        r[:-3] += rng.normal(0.0, rel_sigma * h, r0.size - 3)
        params = dict(kc=units.kc_from_kT(kc, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0, mode=1)
        dt = min(dt_max, dt_safety * dt_max * (h / h_ref) ** 2)
        out.append(Member(L, R, kc, int(seed * 100003 + i), path, r0, r, params, T, dt, h))
    return out


@dataclass
class ServeRequest:
    """One request of the serving stream."""
    kind: str            # "icosphere" or "sheet"
    size: int            # subdivision level, or vertices per side
    mesh_path: str
    r0: np.ndarray       # setg0 state
    r: np.ndarray        # start positions
    params: dict
    c0: float
    steps: int
    seed: int
    dt: float
    h: float
    arrival: float       # seconds after the stream starts
    T: float = 298.0

    @property
    def nv(self) -> int:
        return (self.r.size - 3) // 3

    @property
    def nfaces(self) -> int:
        return 2 * self.nv - (4 if self.kind == "icosphere" else 0)


def _sheet_template(n: int, Lx: float = 1000.0):
    mesh = meshgen.hex_sheet(n, n, Lx)
    path = _write_once(mesh, "sheet_%d" % n)
    return path, mesh


def serving(n: int = 64, rate: float = 200.0, seed: int = 5, levels=(2, 3, 4, 5, 6), sheet_sides=(32, 256),
            steps_range=(100, 10000), kc_range=(10.0, 40.0), c0_range=(-0.01, 0.01), KA_kT: float = 0.215,
            T: float = 298.0, dt_max: float = 1e-9, h_ref: float = 13.0, dt_safety: float = 0.3) -> list[ServeRequest]:
    """The serving stream: Poisson arrivals at `rate` per second; each request
    is, with equal probability, an icosphere of level 2-6 (radius 1000 A
    2^(L-6), so every level has ~16 A edges, isotropic noise 5 % of the edge)
    or a periodic hex sheet of N x N vertices, N even in [32, 256], in a
    1000 x 866 A box (heights: Fourier modes with amplitude ~ q^-2).  Steps
    are log-uniform in steps_range, kc ~ U[kc_range] kT and c0 ~ U[c0_range]
    1/A per request (Helfrich form, local KA = 0.215 kT/A^2), D = 1e-12 m^2/s
    and the edge-length time step of `ensemble`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    ico = {}
    sheets = {}
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        kc = float(rng.uniform(*kc_range))
        c0 = float(rng.uniform(*c0_range))
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        rseed = int(seed * 1000003 + i)
        if rng.random() < 0.5:
            L = int(rng.integers(levels[0], levels[-1] + 1))
            if L not in ico:
                ico[L] = icosphere_template(L)
            path, base, h100 = ico[L]
            R = 1000.0 * 2.0 ** (L - 6)
            h = h100 * R / 100.0
            r0 = base.copy()
            r0[:-3] *= R / 100.0
            r = r0.copy()
            r[:-3] += rng.normal(0.0, 0.05 * h, r0.size - 3)
            kind, size = "icosphere", L
        else:
            N = 2 * int(rng.integers(sheet_sides[0] // 2, sheet_sides[1] // 2 + 1))
            if N not in sheets:
                sheets[N] = _sheet_template(N)
            path, mesh = sheets[N]
            h = mesh.pbc[0, 0] / N
            r0 = meshgen.flat_coords(mesh)
            r = r0.copy()
            r[2:-3:3] += meshgen.fourier_heights(mesh, kc, 6, rseed)
            kind, size = "sheet", N
        params = dict(kc=units.kc_from_kT(kc, T), kg=0.0, KA=units.KA_from_kT(KA_kT, T), k_reg=0.0, mode=1)
        dt = min(dt_max, dt_safety * dt_max * (h / h_ref) ** 2)
        out.append(ServeRequest(kind, size, path, r0, r, params, c0, steps, rseed, dt, h, t, T))
    return out
