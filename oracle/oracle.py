"""ctypes wrapper around the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, as the checker.  It mirrors the reference ``surface`` calls:
``loadLattice`` -> ``generatePlan`` -> ``setg0`` -> ``energy`` / ``grad``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODE_LIPID = 0   # FIXED_A + LOCAL_LIPID_ENERGY (reference default)
MODE_HELFRICH = 1
GRAD_REFERENCE = 0   # surface::grad's Helfrich gradient (src/grad.C:448)
GRAD_EXACT = 1       # the exact derivative of the Helfrich energy


class OParams(C.Structure):
    _fields_ = [("kc", C.c_double), ("kg", C.c_double), ("KA", C.c_double),
                ("k_reg", C.c_double), ("mode", C.c_int), ("lipid_c0", C.c_double),
                ("lipid_apl", C.c_double), ("grad_form", C.c_int)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.orc_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.orc_free.argtypes = [vp]
        L.orc_nv.argtypes = [vp]
        L.orc_nfaces.argtypes = [vp, ip, ip]
        L.orc_get_positions.argtypes = [vp, dp]
        L.orc_plan.argtypes = [vp]
        L.orc_set_params.argtypes = [vp, C.POINTER(OParams)]
        L.orc_set_face_c0.argtypes = [vp, dp]
        L.orc_setg0.argtypes = [vp, dp]
        L.orc_energy.argtypes = [vp, dp]
        L.orc_energy.restype = C.c_double
        L.orc_face_energy.argtypes = [vp, C.c_int, dp]
        L.orc_face_energy.restype = C.c_double
        L.orc_grad.argtypes = [vp, dp, dp]
        L.orc_frozen_curvature_area.argtypes = [vp, dp, dp]
        L.orc_frozen_curvature_area.restype = C.c_double
        L.orc_area.argtypes = [vp, dp, dp, dp]
        L.orc_bd_steps.argtypes = [vp, dp, dp, C.c_int, C.c_double, C.c_double, dp]
        L.orc_face_info.argtypes = [vp, C.c_int, ip, ip, ip, ip]
        L.orc_face_table.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_tri_lipids.argtypes = [vp, C.c_int, dp, dp]
        L.orc_tri_lipid_types.argtypes = [vp, C.c_int, dp, dp]
        L.orc_set_lipids.argtypes = [vp, C.c_int, dp, dp, dp, dp, C.c_uint]
        L.orc_grand_seed.argtypes = [vp, C.c_uint]
        L.orc_grand_next.argtypes = [vp]
        L.orc_ntri.argtypes = [vp]
        L.orc_tri_order_face.argtypes = [vp, C.c_int]
        L.orc_neighbors.argtypes = [vp, C.c_int, ip, ip]
        L.orc_get_v_pts.argtypes = [dp, dp, C.c_int, C.c_double]
        L.orc_get_u_pts.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double]
        L.orc_basis.argtypes = [C.c_double, C.c_double, dp]
        L.orc_irregular_table.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.orc_set_restraints.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_limit_points.argtypes = [vp, dp, dp]
        L.orc_volume.argtypes = [vp, dp]
        L.orc_volume.restype = C.c_double
        L.orc_volume_grad.argtypes = [vp, dp, dp]
        L.orc_restraint_energy.argtypes = [vp, dp, dp, dp]
        L.orc_restraint_energy.restype = C.c_double
        L.orc_energy_grad_threads.argtypes = [vp, dp, dp, C.c_int]
        L.orc_energy_grad_threads.restype = C.c_double
        L.orc_bd_steps_threads.argtypes = [vp, dp, dp, C.c_int, C.c_double, C.c_double, dp, C.c_int]
        L.orc_effm_points_per_face.restype = C.c_int
        L.orc_effm_point_mass.argtypes = [vp]
        L.orc_effm_point_mass.restype = C.c_double
        L.orc_effm_mass.argtypes = [vp, C.c_double, dp]
        L.orc_effm_invert.argtypes = [C.c_int, dp]
        L.orc_effm_invert.restype = C.c_int
        L.orc_effm_truncate.argtypes = [C.c_int, dp, C.c_double]
        L.orc_effm_truncate.restype = C.c_long
        L.orc_effm_steps.argtypes = [vp, dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     dp, dp]
        L.orc_point_eval.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, dp, dp]
        L.orc_point_eval.restype = C.c_int
        L.orc_point_grad.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, dp, dp, C.c_double,
                                     C.c_int, C.c_double, dp, dp]
        L.orc_point_grad.restype = C.c_int
        L.orc_point_geometry.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, dp]
        L.orc_point_geometry.restype = C.c_int
        L.orc_face_density.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_face_density.restype = C.c_int
        L.orc_particle_energy.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, dp]
        L.orc_particle_energy.restype = C.c_int
        L.orc_particle_grad.argtypes = [vp, dp, C.c_int, ip, dp, dp, dp, dp, dp]
        L.orc_particle_grad.restype = C.c_int
        L.orc_particle_curvature_energy.argtypes = [vp, dp, C.c_int, ip, dp, dp, dp]
        L.orc_particle_curvature_energy.restype = C.c_double
        L.orc_particle_density_frozen.argtypes = [vp, dp, dp, C.c_int, ip, dp, dp]
        L.orc_particle_density_frozen.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSurface:
    """The reference surface on the CPU: load, plan, setg0, energy, grad."""

    def __init__(self, mesh_path: str):
        L = lib()
        h = C.c_void_p()
        rc = L.orc_load(mesh_path.encode(), C.byref(h))
        if rc != 0:
            raise ValueError("oracle: cannot load %s (code %d)" % (mesh_path, rc))
        self._h = h
        self.nv = L.orc_nv(h)
        rc = L.orc_plan(h)
        if rc != 0:
            raise ValueError("oracle: illegal mesh for plan (code %d)" % rc)
        nr, ni = C.c_int(), C.c_int()
        L.orc_nfaces(h, C.byref(nr), C.byref(ni))
        self.nreg, self.nirr = nr.value, ni.value
        self.nfaces = self.nreg + self.nirr

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.orc_free(h)
            self._h = None

    def positions(self) -> np.ndarray:
        r = np.zeros(3 * self.nv + 3)
        lib().orc_get_positions(self._h, _p(r))
        return r

    def set_params(self, kc=14.0, kg=0.0, KA=0.0, k_reg=0.0, mode=MODE_LIPID,
                   lipid_c0=0.0, lipid_apl=68.3, grad_form=GRAD_REFERENCE):
        p = OParams(kc, kg, KA, k_reg, mode, lipid_c0, lipid_apl, grad_form)
        lib().orc_set_params(self._h, C.byref(p))

    def frozen_curvature_area(self, r_star: np.ndarray, r: np.ndarray) -> float:
        """sum over regular-face Gauss points of 1/2 kc (H(r_star) - c0)^2 dudv w g(r)"""
        r_star = np.ascontiguousarray(r_star, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        return lib().orc_frozen_curvature_area(self._h, _p(r_star), _p(r))

    def set_face_c0(self, c0: np.ndarray):
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        assert c0.size == self.nfaces
        lib().orc_set_face_c0(self._h, _p(c0))

    def setg0(self, r: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float64)
        lib().orc_setg0(self._h, _p(r))

    def energy(self, r: np.ndarray) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        return lib().orc_energy(self._h, _p(r))

    def face_energy(self, f: int, r: np.ndarray) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        return lib().orc_face_energy(self._h, f, _p(r))

    def grad(self, r: np.ndarray, g: np.ndarray | None = None) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        if g is None:
            g = np.zeros_like(r)
        lib().orc_grad(self._h, _p(r), _p(g))
        return g

    def area(self, r: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float64)
        a, a0 = C.c_double(), C.c_double()
        lib().orc_area(self._h, _p(r), C.byref(a), C.byref(a0))
        return a.value, a0.value

    def set_restraints(self, KA_global: float = 0.0, KV: float = 0.0, A0: float = 0.0,
                       V0: float = 0.0):
        """Global area / volume restraints; A0 <= 0 and V0 == 0 take the
        setg0 reference values."""
        lib().orc_set_restraints(self._h, KA_global, KV, A0, V0)

    def limit_points(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        L = np.zeros(3 * self.nv)
        lib().orc_limit_points(self._h, _p(r), _p(L))
        return L

    def volume(self, r: np.ndarray) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        return lib().orc_volume(self._h, _p(r))

    def volume_grad(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        d = np.zeros(3 * self.nv)
        lib().orc_volume_grad(self._h, _p(r), _p(d))
        return d

    def restraint_energy(self, r: np.ndarray):
        """(E_restraint, dE/dA_tot, dE/dV) at r."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        la, lv = C.c_double(), C.c_double()
        e = lib().orc_restraint_energy(self._h, _p(r), C.byref(la), C.byref(lv))
        return e, la.value, lv.value

    def bd_steps(self, r: np.ndarray, xi: np.ndarray, mobility_dt: float,
                 noise_amp: float, want_energies: bool = False):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        n = xi.size // (3 * self.nv)
        en = np.zeros(n) if want_energies else None
        lib().orc_bd_steps(self._h, _p(r), _p(xi), n, mobility_dt, noise_amp,
                           _p(en) if en is not None else None)
        return (r, en) if want_energies else r

    def bd_steps_threads(self, r: np.ndarray, xi: np.ndarray, mobility_dt: float, noise_amp: float,
                         nthreads: int) -> np.ndarray:
        """bd_steps with each energy + gradient over nthreads OpenMP threads."""
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        n = xi.size // (3 * self.nv)
        lib().orc_bd_steps_threads(self._h, _p(r), _p(xi), n, mobility_dt, noise_amp, None, int(nthreads))
        return r

    def energy_grad_threads(self, r: np.ndarray, nthreads: int):
        """(E, g) with the faces split over nthreads OpenMP threads."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        g = np.zeros_like(r)
        e = lib().orc_energy_grad_threads(self._h, _p(r), _p(g), int(nthreads))
        return e, g

    # effective-mass integrators (src/hd.C, src/modes.C)
    def effm_point_mass(self) -> float:
        return lib().orc_effm_point_mass(self._h)

    def effm_mass(self, mass_per_pt: float | None = None) -> np.ndarray:
        m = self.effm_point_mass() if mass_per_pt is None else mass_per_pt
        M = np.zeros((self.nv, self.nv))
        lib().orc_effm_mass(self._h, m, _p(M))
        return M

    @staticmethod
    def effm_invert(M: np.ndarray) -> np.ndarray:
        A = np.ascontiguousarray(M, dtype=np.float64).copy()
        rc = lib().orc_effm_invert(A.shape[0], _p(A))
        if rc != 0:
            raise ValueError("oracle: mass matrix not positive definite (pivot %d)" % (rc - 1))
        return A

    @staticmethod
    def effm_truncate(Mi: np.ndarray, rel_tol: float = 1e-20):
        A = np.ascontiguousarray(Mi, dtype=np.float64).copy()
        kept = lib().orc_effm_truncate(A.shape[0], _p(A), rel_tol)
        return A, int(kept)

    def effm_inverse(self) -> np.ndarray:
        """The truncated inverse effective mass the integrator uses."""
        return self.effm_truncate(self.effm_invert(self.effm_mass()))[0]

    def effm_steps(self, r: np.ndarray, pp: np.ndarray, xi: np.ndarray, mode: int, dt: float, kT: float,
                   gamma: float, Mi: np.ndarray, energies: bool = False):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        pp = np.ascontiguousarray(pp, dtype=np.float64).copy()
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        Mi = np.ascontiguousarray(Mi, dtype=np.float64)
        n = xi.shape[0]
        e = np.zeros(n)
        lib().orc_effm_steps(self._h, _p(r), _p(pp), _p(xi), n, int(mode), dt, kT, gamma, _p(Mi),
                             _p(e) if energies else None)
        return (r, pp, e) if energies else (r, pp)

    # points on the limit surface (src/interpLib.C:7936, src/grad.C:5697)
    def point_eval(self, r: np.ndarray, faces, uv):
        """(R, nrm), each [n, 3], at the points (faces[i], uv[i])."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        n = uv.shape[0]
        R, N = np.zeros((n, 3)), np.zeros((n, 3))
        for i in range(n):
            a, b = np.zeros(3), np.zeros(3)
            rc = lib().orc_point_eval(self._h, _p(r), int(faces[i]), float(uv[i, 0]), float(uv[i, 1]),
                                      _p(a), _p(b))
            if rc:
                raise ValueError("oracle: point %d outside the reference's domain" % i)
            R[i], N[i] = a, b
        return R, N

    def point_grad(self, r: np.ndarray, faces, uv, de_dr, de_dnrm, frac_mult=None,
                   on_surface: int = 0, kT: float = 0.0, g: np.ndarray | None = None):
        """(g, g_puv): pointGradient of every point added, in point order, to
        g (3nv+3; zeros if None) and to g_puv [n, 2] (zeros)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        de_dr = np.asarray(de_dr, dtype=np.float64).reshape(-1, 3)
        de_dnrm = np.asarray(de_dnrm, dtype=np.float64).reshape(-1, 3)
        n = uv.shape[0]
        g = np.zeros_like(r) if g is None else np.ascontiguousarray(g, dtype=np.float64)
        gp = np.zeros((n, 2))
        for i in range(n):
            a = np.ascontiguousarray(de_dr[i])
            b = np.ascontiguousarray(de_dnrm[i])
            t = np.zeros(2)
            fm = 1.0 if frac_mult is None else float(frac_mult[i])
            rc = lib().orc_point_grad(self._h, _p(r), int(faces[i]), float(uv[i, 0]), float(uv[i, 1]),
                                      _p(a), _p(b), fm, int(on_surface), float(kT), _p(g), _p(t))
            if rc:
                raise ValueError("oracle: point %d outside the reference's domain" % i)
            gp[i] = t
        return g, gp

    def point_geometry(self, r: np.ndarray, faces, uv) -> np.ndarray:
        """[n, 15]: Ru, Rv, g, c1 + c2, c1 c2, c1, c2, vec1, vec2 at the points."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 15))
        for i in range(uv.shape[0]):
            t = np.zeros(15)
            if lib().orc_point_geometry(self._h, _p(r), int(faces[i]), float(uv[i, 0]), float(uv[i, 1]), _p(t)):
                raise ValueError("oracle: point %d outside the reference's domain" % i)
            out[i] = t
        return out

    def face_density(self, r: np.ndarray, f: int):
        """(numerator, area) of the face's curvature-energy density (penergy's
        face_energy_density before the division)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        num, area = C.c_double(), C.c_double()
        if lib().orc_face_density(self._h, int(f), _p(r), C.byref(num), C.byref(area)):
            raise ValueError("oracle: face out of range")
        return num.value, area.value

    @staticmethod
    def _particles(faces, uv, c0, area):
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1)
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        area = np.ascontiguousarray(area, dtype=np.float64)
        return faces, uv, c0, area, faces.ctypes.data_as(C.POINTER(C.c_int))

    def particle_grad(self, r: np.ndarray, faces, uv, c0, area, g=None, g_puv=None):
        """pgrad plus grad's particle-area terms: (g, g_puv), accumulated into
        the given arrays (zeros when None)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        faces, uv, c0, area, fp = self._particles(faces, uv, c0, area)
        n = faces.size
        g = np.zeros_like(r) if g is None else g
        g_puv = np.zeros(2 * n) if g_puv is None else g_puv
        rc = lib().orc_particle_grad(self._h, _p(r), n, fp, _p(uv), _p(c0), _p(area), _p(g), _p(g_puv))
        if rc:
            raise ValueError("oracle: particle %d outside the reference's domain" % (-1 - rc))
        return g, g_puv

    def particle_curvature_energy(self, r: np.ndarray, faces, uv, c0, area) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        faces, uv, c0, area, fp = self._particles(faces, uv, c0, area)
        return lib().orc_particle_curvature_energy(self._h, _p(r), faces.size, fp, _p(uv), _p(c0), _p(area))

    def particle_density_frozen(self, r_star: np.ndarray, r: np.ndarray, faces, c0, area) -> float:
        r_star = np.ascontiguousarray(r_star, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        faces, _, c0, area, fp = self._particles(faces, np.zeros(2 * len(faces)), c0, area)
        return lib().orc_particle_density_frozen(self._h, _p(r_star), _p(r), faces.size, fp, _p(c0), _p(area))

    def particle_energy(self, r: np.ndarray, faces, uv, c0, area) -> np.ndarray:
        """penergy's term of every particle: 1/2 kc a (c1 + c2 - c0)^2 - a rho_f."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        dens = {}
        out = np.zeros(uv.shape[0])
        for i in range(uv.shape[0]):
            f = int(faces[i])
            if f not in dens:
                num, ar = self.face_density(r, f)
                dens[f] = num / ar
            t = C.c_double()
            if lib().orc_particle_energy(self._h, _p(r), f, float(uv[i, 0]), float(uv[i, 1]), float(c0[i]),
                                         float(area[i]), dens[f], C.byref(t)):
                raise ValueError("oracle: particle %d outside the reference's domain" % i)
            out[i] = t.value
        return out

    def face_info(self, f: int):
        owner, ncp, tri = C.c_int(), C.c_int(), C.c_int()
        cp = (C.c_int * 16)()
        npt = lib().orc_face_info(self._h, f, C.byref(owner), C.byref(ncp), cp, C.byref(tri))
        return dict(owner=owner.value, ncp=ncp.value, cp=list(cp)[:ncp.value], tri=tri.value,
                    npt=npt)

    def face_table(self, f: int, p: int):
        w6 = np.zeros(6 * 16)
        pbc = np.zeros(48)
        wt, g0 = C.c_double(), C.c_double()
        ncp = lib().orc_face_table(self._h, f, p, _p(w6), C.byref(wt), _p(pbc), C.byref(g0))
        return (w6[:6 * ncp].reshape(6, ncp), wt.value, pbc[:3 * ncp].reshape(ncp, 3), g0.value)

    def tri_lipids(self, t: int):
        lo, li = C.c_double(), C.c_double()
        lib().orc_tri_lipids(self._h, t, C.byref(lo), C.byref(li))
        return lo.value, li.value

    def set_lipids(self, c0, apl, inner, outer, seed: int = 1):
        """A lipid composition for the next setg0 (types' c0, area per lipid,
        input parts per leaflet; packing with rand() after srand(seed))."""
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (c0, apl, inner, outer)]
        self._ntypes = len(arr[0])
        rc = lib().orc_set_lipids(self._h, self._ntypes, *[_p(x) for x in arr], int(seed))
        if rc != 0:
            raise ValueError("bad lipid composition")

    def tri_lipid_types(self, t: int):
        n = getattr(self, "_ntypes", 1)
        lo, li = np.zeros(n), np.zeros(n)
        lib().orc_tri_lipid_types(self._h, t, _p(lo), _p(li))
        return lo, li

    def ntri(self) -> int:
        return lib().orc_ntri(self._h)

    def tri_face(self, t: int) -> int:
        return lib().orc_tri_order_face(self._h, t)

    def neighbors(self, v: int):
        val = C.c_int()
        nb = (C.c_int * 16)()
        lib().orc_neighbors(self._h, v, C.byref(val), nb)
        return list(nb)[:val.value]


def v_points(k: float, n: int = 4):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_get_v_pts(_p(x), _p(w), n, k)
    return x, w


def u_points(k: float, v: float, n: int = 4):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_get_u_pts(_p(x), _p(w), n, k, v)
    return x, w


def basis(u: float, v: float) -> np.ndarray:
    out = np.zeros(72)
    lib().orc_basis(u, v, _p(out))
    return out.reshape(6, 12)


def irregular_table(val: int, u: float, v: float) -> np.ndarray:
    out = np.zeros(6 * (val + 6))
    rc = lib().orc_irregular_table(val, u, v, _p(out))
    if rc:
        raise ValueError("bad valence")
    return out.reshape(6, val + 6)


def glibc_rand(seed: int, n: int) -> np.ndarray:
    """The oracle's restatement of glibc rand(): n outputs after srand(seed)."""
    st = (C.c_int32 * 33)()
    lib().orc_grand_seed(st, int(seed))
    return np.array([lib().orc_grand_next(st) for _ in range(n)], dtype=np.int64)

