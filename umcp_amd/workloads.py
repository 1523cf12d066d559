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
  N(0, 2^2) noise, kc = 20 kT, per-face c0 ~ U[-0.01, 0.01] 1/A (Helfrich form).

Reference state: the unperturbed mesh (r0) is the setg0 state; the energy is
evaluated at the perturbed r.
"""
from __future__ import annotations

import hashlib
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


def vesicle_642(seed: int = 1) -> Workload:
    return vesicle(3, 100.0, 0.5, seed, name="vesicle_642")


def vesicle_40962(seed: int = 3, with_c0: bool = True) -> Workload:
    w = vesicle(6, 1000.0, 2.0, seed, radial=True, name="vesicle_40962")
    if with_c0:
        w.params["mode"] = 1
        w.extra["c0_seed"] = seed + 100
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
