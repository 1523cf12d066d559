"""Python mirror of the reference ``surface`` interface over libumcp_core.so.

The reference drives its hot path through C++ methods of ``struct surface``
(src/interp.h:210-552): ``loadLattice`` (291), ``generatePlan`` (337),
``setg0`` (293), ``energy`` (294), ``grad`` (304, accumulating), ``area``
(301).  ``Surface`` keeps those names, argument meanings and the
``r``/``g`` layout (3 nv coordinates + the alpha triple), and calls the HIP
library through its C ABI (include/umcp.h).  There is no CPU fallback: if the
library or a GPU is missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import units

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libumcp_core.so")

MODE_LIPID = 0
MODE_HELFRICH = 1
PREC_FP64 = 0
PREC_MIXED = 1
GRAD_REFERENCE = 0   # surface::grad's Helfrich gradient (src/grad.C:448)
GRAD_EXACT = 1       # the exact derivative of the Helfrich energy
LAUNCH_OCC2 = 2
LAUNCH_MFMA = 4
LAUNCH_CPS = 8

# every symbol include/umcp.h declares
EXPORTS = [
    "umcp_link_check", "umcp_set_device", "umcp_load_mesh", "umcp_destroy", "umcp_last_error",
    "umcp_nv", "umcp_nfaces", "umcp_get", "umcp_set_params", "umcp_prepare", "umcp_plan", "umcp_set_face_c0",
    "umcp_setg0", "umcp_energy", "umcp_grad", "umcp_energy_grad", "umcp_area", "umcp_bd_steps",
    "umcp_energy_grad_dev", "umcp_energy_grad_dev_noalloc", "umcp_bd_steps_dev", "umcp_class_count", "umcp_class_info",
    "umcp_class_table", "umcp_face_points", "umcp_face_lipids", "umcp_set_launch",
    "umcp_get_launch", "umcp_profile_begin", "umcp_profile_end", "umcp_profile_sample",
    "umcp_set_restraints", "umcp_volume", "umcp_set_step", "umcp_get_step", "umcp_bd_steps_rng_dev",
    "umcp_noise", "umcp_ensemble_create", "umcp_ensemble_destroy", "umcp_ensemble_last_error",
    "umcp_ensemble_add", "umcp_ensemble_remove", "umcp_ensemble_size", "umcp_ensemble_refresh",
    "umcp_ensemble_set_launch", "umcp_ensemble_get_launch", "umcp_ensemble_bd_steps", "umcp_ensemble_work",
    "umcp_nan_check", "umcp_ensemble_nan_check", "umcp_read_lipid_composition", "umcp_set_lipids",
    "umcp_lipid_type_count", "umcp_face_lipid_types",
    "umcp_clone", "umcp_device_bytes", "umcp_server_create", "umcp_server_submit", "umcp_server_wait",
    "umcp_server_drain", "umcp_server_stats", "umcp_server_last_error", "umcp_server_destroy",
    "umcp_effm_build", "umcp_effm_get", "umcp_effm_steps", "umcp_effm_steps_dev",
    "umcp_effm_build_sparse", "umcp_effm_apply", "umcp_point_eval_dev", "umcp_point_status", "umcp_point_grad_dev", "umcp_point_geometry_dev",
    "umcp_save_mesh", "umcp_write_restart", "umcp_read_restart", "umcp_write_limit_surface",
    "umcp_point_eval", "umcp_point_grad", "umcp_point_geometry", "umcp_write_limit_psf",
    "umcp_face_energies", "umcp_energy_vertex", "umcp_area_vertex", "umcp_next_face", "umcp_point_coeffs",
    "umcp_particle_energy", "umcp_particle_grad",
]


class UmcpError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("kc", C.c_double), ("kg", C.c_double), ("KA", C.c_double), ("k_reg", C.c_double),
                ("mode", C.c_int), ("precision", C.c_int), ("lipid_c0", C.c_double),
                ("lipid_apl", C.c_double), ("grad_form", C.c_int)]


class BDParams(C.Structure):
    _fields_ = [("mobility_dt", C.c_double), ("noise_amp", C.c_double)]


EFFM_BD = 0
EFFM_LD = 1
EFFM_MAX_NV = 65536


class EffmParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("dt", C.c_double), ("kT", C.c_double), ("gamma", C.c_double)]


class Restraints(C.Structure):
    _fields_ = [("KA_global", C.c_double), ("KV", C.c_double), ("A0", C.c_double), ("V0", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_double), ("max_active", C.c_int), ("chunk_steps", C.c_int),
                ("precision", C.c_int), ("point_split", C.c_int), ("flags", C.c_int), ("face_grid", C.c_int),
                ("order", C.c_int), ("graph_steps", C.c_int), ("prep_threads", C.c_int)]


class Request(C.Structure):
    _fields_ = [("mesh_path", C.c_char_p), ("nv", C.c_int), ("params", Params), ("c0", C.c_double),
                ("steps", C.c_int), ("seed", C.c_ulonglong), ("bd", BDParams),
                ("r0", C.POINTER(C.c_double)), ("r", C.POINTER(C.c_double))]


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libumcp_core.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise UmcpError("libumcp_core.so is not built (run python -m umcp_amd.build)")
    L = C.CDLL(path)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    sig = {
        "umcp_link_check": ([], C.c_int),
        "umcp_set_device": ([C.c_int], C.c_int),
        "umcp_load_mesh": ([C.c_char_p, C.POINTER(vp)], C.c_int),
        "umcp_destroy": ([vp], None),
        "umcp_last_error": ([vp], C.c_char_p),
        "umcp_nv": ([vp], C.c_int),
        "umcp_nfaces": ([vp, ip, ip], C.c_int),
        "umcp_get": ([vp, dp], C.c_int),
        "umcp_set_params": ([vp, C.POINTER(Params)], C.c_int),
        "umcp_prepare": ([vp], C.c_int),
        "umcp_plan": ([vp], C.c_int),
        "umcp_set_face_c0": ([vp, dp], C.c_int),
        "umcp_setg0": ([vp, dp], C.c_int),
        "umcp_energy": ([vp, dp, dp], C.c_int),
        "umcp_grad": ([vp, dp, dp], C.c_int),
        "umcp_energy_grad": ([vp, dp, dp, dp], C.c_int),
        "umcp_area": ([vp, dp, dp, dp], C.c_int),
        "umcp_bd_steps": ([vp, dp, dp, C.c_int, C.POINTER(BDParams), dp], C.c_int),
        "umcp_energy_grad_dev": ([vp, vp, vp, vp, vp], C.c_int),
        "umcp_energy_grad_dev_noalloc": ([vp, vp, vp, vp, vp], C.c_int),
        "umcp_bd_steps_dev": ([vp, vp, vp, C.c_int, C.POINTER(BDParams), vp, vp], C.c_int),
        "umcp_class_count": ([vp], C.c_int),
        "umcp_class_info": ([vp, C.c_int, ip, ip, ip, ip], C.c_int),
        "umcp_class_table": ([vp, C.c_int, dp, dp], C.c_int),
        "umcp_face_points": ([vp, C.c_int, ip, ip], C.c_int),
        "umcp_face_lipids": ([vp, C.c_int, dp, dp], C.c_int),
        "umcp_set_launch": ([vp, C.c_int, C.c_int], C.c_int),
        "umcp_get_launch": ([vp, ip, ip], C.c_int),
        "umcp_profile_begin": ([vp], C.c_int),
        "umcp_profile_sample": ([vp, C.c_int], C.c_int),
        "umcp_profile_end": ([vp, C.c_int, dp, ip], C.c_int),
        "umcp_set_restraints": ([vp, C.POINTER(Restraints)], C.c_int),
        "umcp_volume": ([vp, dp, dp], C.c_int),
        "umcp_set_step": ([vp, C.c_int], C.c_int),
        "umcp_get_step": ([vp, ip], C.c_int),
        "umcp_bd_steps_rng_dev": ([vp, vp, C.c_int, C.POINTER(BDParams), C.c_ulonglong, vp, C.c_int, vp],
                                  C.c_int),
        "umcp_noise": ([C.c_ulonglong, C.c_int, C.c_int, vp, vp], C.c_int),
        "umcp_save_mesh": ([vp, dp, C.c_char_p], C.c_int),
        "umcp_write_limit_surface": ([vp, dp, C.c_char_p, C.c_int], C.c_int),
        "umcp_point_eval": ([vp, dp, C.c_int, ip, dp, dp, dp], C.c_int),
        "umcp_write_limit_psf": ([vp, C.c_char_p], C.c_int),
        "umcp_point_geometry": ([vp, dp, C.c_int, ip, dp, dp], C.c_int),
        "umcp_face_energies": ([vp, dp, dp, dp, dp], C.c_int),
        "umcp_energy_vertex": ([vp, dp, C.c_int, dp], C.c_int),
        "umcp_area_vertex": ([vp, dp, C.c_int, dp, dp], C.c_int),
        "umcp_next_face": ([vp, dp, C.c_int, ip, dp, dp, dp, dp], C.c_int),
        "umcp_point_coeffs": ([vp, C.c_int, ip, dp, ip, dp], C.c_int),
        "umcp_particle_energy": ([vp, dp, C.c_int, ip, dp, dp, dp, dp], C.c_int),
        "umcp_particle_grad": ([vp, dp, C.c_int, ip, dp, dp, dp, dp, dp], C.c_int),
        "umcp_point_grad": ([vp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, C.c_double, dp, dp], C.c_int),
        "umcp_write_restart": ([C.c_char_p, C.c_int, dp, dp, C.c_int], C.c_int),
        "umcp_read_restart": ([C.c_char_p, C.c_int, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "umcp_effm_build": ([vp, C.POINTER(C.c_long), dp], C.c_int),
        "umcp_effm_get": ([vp, dp], C.c_int),
        "umcp_effm_build_sparse": ([vp, C.c_double, C.POINTER(C.c_long), dp, C.POINTER(C.c_int)], C.c_int),
        "umcp_effm_apply": ([vp, dp, dp], C.c_int),
        "umcp_point_eval_dev": ([vp, vp, C.c_int, vp, vp, vp, vp, vp], C.c_int),
        "umcp_point_status": ([vp, C.POINTER(C.c_int)], C.c_int),
        "umcp_point_geometry_dev": ([vp, vp, C.c_int, vp, vp, vp, vp], C.c_int),
        "umcp_point_grad_dev": ([vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_double, vp, vp, vp], C.c_int),
        "umcp_effm_steps": ([vp, dp, dp, dp, C.c_int, C.POINTER(EffmParams), dp], C.c_int),
        "umcp_effm_steps_dev": ([vp, vp, vp, vp, C.c_int, C.POINTER(EffmParams), C.c_ulonglong, vp, vp],
                                C.c_int),
        "umcp_ensemble_create": ([C.POINTER(vp)], C.c_int),
        "umcp_ensemble_destroy": ([vp], None),
        "umcp_ensemble_last_error": ([vp], C.c_char_p),
        "umcp_ensemble_add": ([vp, vp, vp, C.POINTER(BDParams), C.c_ulonglong, vp, C.c_int], C.c_int),
        "umcp_ensemble_remove": ([vp, vp], C.c_int),
        "umcp_ensemble_size": ([vp], C.c_int),
        "umcp_ensemble_refresh": ([vp], C.c_int),
        "umcp_ensemble_set_launch": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "umcp_ensemble_get_launch": ([vp, ip, ip, ip, ip], C.c_int),
        "umcp_ensemble_bd_steps": ([vp, C.c_int, C.c_int, vp], C.c_int),
        "umcp_ensemble_work": ([vp, ip, ip, ip], C.c_int),
        "umcp_nan_check": ([vp], C.c_int),
        "umcp_read_lipid_composition": ([vp, C.c_char_p, C.c_uint], C.c_int),
        "umcp_set_lipids": ([vp, C.c_int, dp, dp, dp, dp, C.c_uint], C.c_int),
        "umcp_lipid_type_count": ([vp], C.c_int),
        "umcp_face_lipid_types": ([vp, C.c_int, dp, dp], C.c_int),
        "umcp_ensemble_nan_check": ([vp, ip], C.c_int),
        "umcp_clone": ([vp, C.POINTER(vp)], C.c_int),
        "umcp_device_bytes": ([vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
        "umcp_server_create": ([C.POINTER(ServerConfig), C.POINTER(vp)], C.c_int),
        "umcp_server_submit": ([vp, C.POINTER(Request), ip], C.c_int),
        "umcp_server_wait": ([vp, C.c_int, dp, dp, dp], C.c_int),
        "umcp_server_drain": ([vp], C.c_int),
        "umcp_server_stats": ([vp, dp, C.c_int], C.c_int),
        "umcp_server_last_error": ([vp], C.c_char_p),
        "umcp_server_destroy": ([vp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Surface:
    """One membrane surface resident on a GPU.

    Mirrors the reference call sequence of src/min.C:68-99:
    ``loadLattice(path)`` -> ``generatePlan()`` -> ``get()`` -> ``setg0(r)``
    -> ``energy(r)`` / ``grad(r, g)``.
    """

    def __init__(self, mesh_path: str | None = None, device: int = 0):
        self._L = load_library()
        self._h = None
        self.device = device
        if mesh_path is not None:
            self.loadLattice(mesh_path)

    # -- lifecycle -------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.umcp_last_error(self._h).decode() if self._h else ""
            raise UmcpError("%s failed (%d): %s" % (what, rc, msg))

    def loadLattice(self, path: str, noise: float = 0.0):
        if noise != 0.0:
            raise UmcpError("loadLattice noise is not supported; perturb r instead")
        self._check(self._L.umcp_set_device(self.device), "umcp_set_device")
        h = C.c_void_p()
        rc = self._L.umcp_load_mesh(path.encode(), C.byref(h))
        self._h = h if h.value else None
        self._check(rc, "loadLattice")
        self.nv = self._L.umcp_nv(self._h)
        return 0

    def close(self):
        if self._h is not None:
            self._L.umcp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -----------------------------------------------------------
    def set_params(self, kc=14.0, kg=0.0, KA=0.0, k_reg=0.0, mode=MODE_LIPID,
                   precision=PREC_FP64, lipid_c0=0.0, lipid_apl=68.3, grad_form=GRAD_REFERENCE):
        p = Params(kc, kg, KA, k_reg, mode, precision, lipid_c0, lipid_apl, grad_form)
        self._check(self._L.umcp_set_params(self._h, C.byref(p)), "set_params")

    def generatePlan(self):
        self._check(self._L.umcp_plan(self._h), "generatePlan")
        nr, ni = C.c_int(), C.c_int()
        self.nfaces = self._L.umcp_nfaces(self._h, C.byref(nr), C.byref(ni))
        self.nreg, self.nirr = nr.value, ni.value

    def get(self) -> np.ndarray:
        r = np.zeros(3 * self.nv + 3)
        self._check(self._L.umcp_get(self._h, _dp(r)), "get")
        return r

    def setg0(self, r: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._L.umcp_setg0(self._h, _dp(r)), "setg0")

    def set_face_c0(self, c0: np.ndarray):
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        if c0.size != self.nfaces:
            raise UmcpError("c0 needs one value per face")
        self._check(self._L.umcp_set_face_c0(self._h, _dp(c0)), "set_face_c0")

    # -- hot path --------------------------------------------------------
    def energy(self, r: np.ndarray) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        E = C.c_double()
        self._check(self._L.umcp_energy(self._h, _dp(r), C.byref(E)), "energy")
        return E.value

    def grad(self, r: np.ndarray, g: np.ndarray) -> np.ndarray:
        """Adds dE/dr (and dE/dalpha) into g, as surface::grad does."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        if g.dtype != np.float64 or not g.flags.c_contiguous:
            raise UmcpError("g must be a contiguous float64 array")
        self._check(self._L.umcp_grad(self._h, _dp(r), _dp(g)), "grad")
        return g

    def energy_grad(self, r: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float64)
        g = np.zeros_like(r)
        E = C.c_double()
        self._check(self._L.umcp_energy_grad(self._h, _dp(r), C.byref(E), _dp(g)), "energy_grad")
        return E.value, g

    def area(self, r: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float64)
        A, A0 = C.c_double(), C.c_double()
        self._check(self._L.umcp_area(self._h, _dp(r), C.byref(A), C.byref(A0)), "area")
        return A.value, A0.value

    def bd_steps(self, r: np.ndarray, xi: np.ndarray, mobility_dt: float, noise_amp: float,
                 want_energies: bool = False):
        """Brownian steps from host arrays; returns the new r (and energies)."""
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        n = xi.size // (3 * self.nv)
        en = np.zeros(max(n, 1)) if want_energies else None
        bd = BDParams(mobility_dt, noise_amp)
        rc = self._L.umcp_bd_steps(self._h, _dp(r), _dp(xi), n, C.byref(bd),
                                   _dp(en) if en is not None else None)
        self._check(rc, "bd_steps")
        return (r, en[:n]) if want_energies else r

    # -- device-resident (torch tensors on the GPU) ----------------------
    def energy_grad_dev(self, d_r, d_E, d_g, stream=None):
        rc = self._L.umcp_energy_grad_dev(self._h, C.c_void_p(d_r.data_ptr()),
                                          C.c_void_p(d_E.data_ptr()),
                                          C.c_void_p(d_g.data_ptr()) if d_g is not None else None,
                                          C.c_void_p(stream) if stream else None)
        self._check(rc, "energy_grad_dev")

    def readLipidComposition(self, path: str, seed: int = 1):
        """surface::readLipidComposition: the composition of a file's "lipid"
        directives, packed at the next setg0 with rand() after srand(seed)."""
        self._check(self._L.umcp_read_lipid_composition(self._h, path.encode(), int(seed)), "readLipidComposition")

    def set_lipids(self, c0, apl, inner, outer, seed: int = 1):
        """A lipid composition from arrays (types' c0 1/A, area per lipid A^2,
        input parts of the inner and outer leaflet)."""
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (c0, apl, inner, outer)]
        self._check(self._L.umcp_set_lipids(self._h, len(arr[0]), *[_dp(x) for x in arr], int(seed)), "set_lipids")

    def face_lipid_types(self, f: int):
        """Each lipid type's area in the outer and inner leaflet of face f."""
        n = self._L.umcp_lipid_type_count(self._h)
        lo, li = np.zeros(n), np.zeros(n)
        self._check(self._L.umcp_face_lipid_types(self._h, int(f), _dp(lo), _dp(li)), "face_lipid_types")
        return lo, li

    def nan_check(self):
        """Raises UmcpError (UMCP_ERR_NAN_ENERGY) if an evaluation since the last
        report had a non-finite energy (device entry points flag it)."""
        self._check(self._L.umcp_nan_check(self._h), "nan_check")

    def energy_grad_dev_noalloc(self, d_r, d_E, d_g, stream=None):
        """energy_grad_dev with the library's static scratch array as work space
        (no device allocation)"""
        rc = self._L.umcp_energy_grad_dev_noalloc(self._h, C.c_void_p(d_r.data_ptr()),
                                                  C.c_void_p(d_E.data_ptr()),
                                                  C.c_void_p(d_g.data_ptr()) if d_g is not None else None,
                                                  C.c_void_p(stream) if stream else None)
        self._check(rc, "energy_grad_dev_noalloc")

    def bd_steps_dev(self, d_r, d_xi, nsteps: int, mobility_dt: float, noise_amp: float,
                     d_energies=None, stream=None):
        bd = BDParams(mobility_dt, noise_amp)
        rc = self._L.umcp_bd_steps_dev(self._h, C.c_void_p(d_r.data_ptr()),
                                       C.c_void_p(d_xi.data_ptr()), nsteps, C.byref(bd),
                                       C.c_void_p(d_energies.data_ptr()) if d_energies is not None
                                       else None, C.c_void_p(stream) if stream else None)
        self._check(rc, "bd_steps_dev")

    def bd_steps_rng_dev(self, d_r, nsteps: int, mobility_dt: float, noise_amp: float, seed: int,
                         d_energies=None, stream=None):
        """Brownian steps with the device noise stream of (seed, step)."""
        bd = BDParams(mobility_dt, noise_amp)
        cap = int(d_energies.numel()) if d_energies is not None else 0
        rc = self._L.umcp_bd_steps_rng_dev(self._h, C.c_void_p(d_r.data_ptr()), nsteps, C.byref(bd),
                                           C.c_ulonglong(seed),
                                           C.c_void_p(d_energies.data_ptr()) if d_energies is not None
                                           else None, cap, C.c_void_p(stream) if stream else None)
        self._check(rc, "bd_steps_rng_dev")

    def save_mesh(self, path: str, r: np.ndarray | None = None):
        """Writes the surface as a .mesh file (positions r, or the file's)."""
        rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._L.umcp_save_mesh(self._h, _dp(rr) if rr is not None else None, path.encode()),
                    "save_mesh")

    def write_limit_surface(self, path: str, r: np.ndarray | None = None, append: bool = False):
        """One XYZ frame of the limit surface at the face corners (writeLimitingSurface)."""
        rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._L.umcp_write_limit_surface(self._h, _dp(rr) if rr is not None else None, path.encode(),
                                                     1 if append else 0), "write_limit_surface")

    # points on the limit surface (src/interpLib.C:7936, src/grad.C:5697)
    def point_eval(self, r: np.ndarray, faces, uv):
        """(R, nrm), each [n, 3]: evaluateRNRM at the points (faces[i], uv[i])."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        n = f.size
        R, N = np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.umcp_point_eval(self._h, _dp(r), n, _ip(f), _dp(uv), _dp(R), _dp(N)), "point_eval")
        return R, N

    def point_eval_dev(self, d_r, d_faces, d_uv, d_R, d_nrm, stream=None):
        """umcp_point_eval on device tensors (d_faces int32 [n], d_uv [n, 2],
        d_R and d_nrm [n, 3] float64); asynchronous: point_status after a
        synchronisation reports a rejected point."""
        n = d_faces.numel()
        rc = self._L.umcp_point_eval_dev(self._h, C.c_void_p(d_r.data_ptr()), n, C.c_void_p(d_faces.data_ptr()),
                                         C.c_void_p(d_uv.data_ptr()), C.c_void_p(d_R.data_ptr()),
                                         C.c_void_p(d_nrm.data_ptr()), C.c_void_p(stream) if stream else None)
        self._check(rc, "point_eval_dev")

    def point_grad_dev(self, d_r, d_faces, d_uv, d_de_dr, d_de_dnrm, d_g, d_g_puv, d_frac=None,
                       on_surface: bool = False, kT: float = 0.0, stream=None):
        """umcp_point_grad on device tensors; d_g (3nv+3) and d_g_puv [n, 2]
        accumulate; asynchronous (point_status after a synchronisation)."""
        n = d_faces.numel()
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        rc = self._L.umcp_point_grad_dev(self._h, ptr(d_r), n, ptr(d_faces), ptr(d_uv), ptr(d_de_dr),
                                         ptr(d_de_dnrm), ptr(d_frac), 1 if on_surface else 0, float(kT),
                                         ptr(d_g), ptr(d_g_puv), C.c_void_p(stream) if stream else None)
        self._check(rc, "point_grad_dev")

    def point_geometry_dev(self, d_r, d_faces, d_uv, d_geom, stream=None):
        """umcp_point_geometry on device tensors (d_geom [n, 15] float64)."""
        n = d_faces.numel()
        rc = self._L.umcp_point_geometry_dev(self._h, C.c_void_p(d_r.data_ptr()), n, C.c_void_p(d_faces.data_ptr()),
                                             C.c_void_p(d_uv.data_ptr()), C.c_void_p(d_geom.data_ptr()),
                                             C.c_void_p(stream) if stream else None)
        self._check(rc, "point_geometry_dev")

    def point_status(self):
        """Raises UmcpError naming the first rejected point of the last
        point_eval_dev (call after synchronising)."""
        bad = C.c_int(-1)
        self._check(self._L.umcp_point_status(self._h, C.byref(bad)), "point_status")

    def point_grad(self, r: np.ndarray, faces, uv, de_dr, de_dnrm, frac_mult=None, on_surface: bool = False,
                   kT: float = 0.0, g: np.ndarray | None = None, g_puv: np.ndarray | None = None):
        """(g, g_puv): pointGradient of every point added to g (3nv+3) and
        g_puv [n, 2] (zeros when not given)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        n = f.size
        a = np.ascontiguousarray(de_dr, dtype=np.float64).reshape(n, 3)
        b = np.ascontiguousarray(de_dnrm, dtype=np.float64).reshape(n, 3)
        fm = None if frac_mult is None else np.ascontiguousarray(frac_mult, dtype=np.float64).reshape(n)
        g = np.zeros_like(r) if g is None else g
        gp = np.zeros((n, 2)) if g_puv is None else g_puv
        assert g.dtype == np.float64 and g.flags.c_contiguous and g.size == r.size
        assert gp.dtype == np.float64 and gp.flags.c_contiguous and gp.size == 2 * n
        self._check(self._L.umcp_point_grad(self._h, _dp(r), n, _ip(f), _dp(uv), _dp(a), _dp(b),
                                            _dp(fm) if fm is not None else None, 1 if on_surface else 0,
                                            float(kT), _dp(g), _dp(gp)), "point_grad")
        return g, gp

    def face_energies(self, r: np.ndarray, areas: bool = False):
        """faceEnergy of every face, in the reference's face order; with
        areas=True also each face's area and reference area: (E, A, A0)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        E = np.zeros(self.nfaces)
        A = np.zeros(self.nfaces) if areas else None
        A0 = np.zeros(self.nfaces) if areas else None
        self._check(self._L.umcp_face_energies(self._h, _dp(r), _dp(E), _dp(A) if areas else None,
                                               _dp(A0) if areas else None), "face_energies")
        return (E, A, A0) if areas else E

    def energy_vertex(self, r: np.ndarray, vertex: int) -> float:
        """energy(r, NULL, do_vertex): the faces the vertex moves, no global terms."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        E = C.c_double()
        self._check(self._L.umcp_energy_vertex(self._h, _dp(r), int(vertex), C.byref(E)), "energy_vertex")
        return E.value

    def particle_energy(self, r: np.ndarray, faces, uv, c0, area) -> np.ndarray:
        """penergy's term of every particle: 1/2 kc a (c1 + c2 - c0)^2 - a rho_f."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        a = np.ascontiguousarray(area, dtype=np.float64)
        assert c0.size == f.size and a.size == f.size
        out = np.zeros(f.size)
        self._check(self._L.umcp_particle_energy(self._h, _dp(r), f.size, _ip(f), _dp(uv), _dp(c0), _dp(a),
                                                 _dp(out)), "particle_energy")
        return out

    def particle_grad(self, r: np.ndarray, faces, uv, c0, area, g=None, g_puv=None):
        """The particle terms of surface::grad (pgrad and the particle-area
        term of faces carrying particles), accumulated into g (3nv+3) and g_puv
        (2n); returns (g, g_puv), zeros first when not given."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        a = np.ascontiguousarray(area, dtype=np.float64)
        assert c0.size == f.size and a.size == f.size
        g = np.zeros(r.size) if g is None else g
        g_puv = np.zeros(2 * f.size) if g_puv is None else g_puv
        assert g.flags.c_contiguous and g_puv.flags.c_contiguous
        self._check(self._L.umcp_particle_grad(self._h, _dp(r), f.size, _ip(f), _dp(uv), _dp(c0), _dp(a), _dp(g),
                                               _dp(g_puv)), "particle_grad")
        return g, g_puv

    def point_coeffs(self, faces, uv):
        """(ids [n, 18], weights [n, 3, 18]): get_pt_coeffs / get_pt_dcoeffs."""
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        ids = np.zeros((f.size, 18), dtype=np.int32)
        w = np.zeros((f.size, 3, 18))
        self._check(self._L.umcp_point_coeffs(self._h, f.size, _ip(f), _dp(uv), _ip(ids), _dp(w)), "point_coeffs")
        return ids, w

    def next_face(self, r: np.ndarray, faces, uv, duv, mom=None, coord_transform=None):
        """One nextFace step per point: returns (faces, uv, duv, mom, coord_transform)
        updated (copies of the inputs)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.array(faces, dtype=np.int32)
        u = np.array(uv, dtype=np.float64).reshape(-1, 2)
        d = np.array(duv, dtype=np.float64).reshape(-1, 2)
        m = None if mom is None else np.array(mom, dtype=np.float64).reshape(-1, 2)
        ct = None if coord_transform is None else np.array(coord_transform, dtype=np.float64).reshape(-1, 4)
        self._check(self._L.umcp_next_face(self._h, _dp(r), f.size, _ip(f), _dp(u), _dp(d),
                                           _dp(m) if m is not None else None,
                                           _dp(ct) if ct is not None else None), "next_face")
        return f, u, d, m, ct

    def area_vertex(self, r: np.ndarray, vertex: int):
        """area(r, do_vertex): (A, A0) of the faces the vertex moves."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        A, A0 = C.c_double(), C.c_double()
        self._check(self._L.umcp_area_vertex(self._h, _dp(r), int(vertex), C.byref(A), C.byref(A0)), "area_vertex")
        return A.value, A0.value

    def point_geometry(self, r: np.ndarray, faces, uv) -> np.ndarray:
        """[n, 15]: Ru, Rv, g, c1 + c2, c1 c2, c1, c2, vec1, vec2 (ru, rv, g, c)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((f.size, 15))
        self._check(self._L.umcp_point_geometry(self._h, _dp(r), f.size, _ip(f), _dp(uv), _dp(out)),
                    "point_geometry")
        return out

    def write_limit_psf(self, path: str):
        """The PSF topology of the limit-surface frames (writeLimitingSurfacePSF)."""
        self._check(self._L.umcp_write_limit_psf(self._h, path.encode()), "write_limit_psf")

    # momentum-space integrators with the effective-mass matrix (src/hd.C)
    def effm_build(self):
        """Builds the truncated inverse effective mass; returns (nnz, point mass)."""
        nnz, m = C.c_long(), C.c_double()
        self._check(self._L.umcp_effm_build(self._h, C.byref(nnz), C.byref(m)), "effm_build")
        return nnz.value, m.value

    def effm_build_sparse(self, tol: float = 0.0):
        """The scalable effective mass: M kept sparse, products by a Chebyshev
        solve to relative residual tol (0: 1e-12).  Returns (nnz of M, point
        mass, iterations per solve)."""
        nnz, m, it = C.c_long(), C.c_double(), C.c_int()
        self._check(self._L.umcp_effm_build_sparse(self._h, float(tol), C.byref(nnz), C.byref(m), C.byref(it)),
                    "effm_build_sparse")
        return nnz.value, m.value, it.value

    def effm_apply(self, v: np.ndarray) -> np.ndarray:
        """Mi v for v of shape (nv, 3) or (3 nv,) with the built form."""
        x = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if x.size != 3 * self.nv:
            raise ValueError("effm_apply: need 3 nv values")
        out = np.empty_like(x)
        self._check(self._L.umcp_effm_apply(self._h, _dp(x), _dp(out)), "effm_apply")
        return out.reshape(np.shape(v))

    def effm_get(self) -> np.ndarray:
        Mi = np.zeros((self.nv, self.nv))
        self._check(self._L.umcp_effm_get(self._h, _dp(Mi)), "effm_get")
        return Mi

    def effm_steps(self, r: np.ndarray, pp: np.ndarray, xi: np.ndarray, mode: int, dt: float, kT: float,
                   gamma: float, energies: bool = False):
        """Host arrays: returns (r, pp) or (r, pp, energies) after xi.shape[0] steps."""
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        pp = np.ascontiguousarray(pp, dtype=np.float64).copy()
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        n = xi.shape[0]
        e = np.zeros(max(n, 1))
        prm = EffmParams(mode, dt, kT, gamma)
        self._check(self._L.umcp_effm_steps(self._h, _dp(r), _dp(pp), _dp(xi), n, C.byref(prm),
                                            _dp(e) if energies else None), "effm_steps")
        return (r, pp, e[:n]) if energies else (r, pp)

    def effm_steps_dev(self, d_r, d_pp, nsteps: int, mode: int, dt: float, kT: float, gamma: float,
                       d_xi=None, seed: int = 0, d_energies=None, stream=None):
        """Device arrays (torch tensors); d_xi None: the device noise stream of (seed, step)."""
        prm = EffmParams(mode, dt, kT, gamma)
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        self._check(self._L.umcp_effm_steps_dev(self._h, ptr(d_r), ptr(d_pp), ptr(d_xi), int(nsteps), C.byref(prm),
                                                C.c_ulonglong(seed), ptr(d_energies),
                                                C.c_void_p(stream) if stream else None), "effm_steps_dev")

    def set_step(self, step: int):
        self._check(self._L.umcp_set_step(self._h, int(step)), "set_step")

    def get_step(self) -> int:
        k = C.c_int()
        self._check(self._L.umcp_get_step(self._h, C.byref(k)), "get_step")
        return k.value

    # -- global restraints -------------------------------------------------
    def set_restraints(self, KA_global: float = 0.0, KV: float = 0.0, A0: float = 0.0, V0: float = 0.0):
        """Global area (KA_global A0 s^2) and volume (KV/2 V0 s^2) restraints;
        A0 <= 0 / V0 == 0 take the setg0 reference values."""
        rs = Restraints(KA_global, KV, A0, V0)
        self._check(self._L.umcp_set_restraints(self._h, C.byref(rs)), "set_restraints")

    def volume(self, r: np.ndarray) -> float:
        r = np.ascontiguousarray(r, dtype=np.float64)
        V = C.c_double()
        self._check(self._L.umcp_volume(self._h, _dp(r), C.byref(V)), "volume")
        return V.value

    # -- introspection ---------------------------------------------------
    def classes(self):
        out = []
        for c in range(self._L.umcp_class_count(self._h)):
            val, ncp, npt, nf = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            self._L.umcp_class_info(self._h, c, C.byref(val), C.byref(ncp), C.byref(npt), C.byref(nf))
            tab = np.zeros(npt.value * 6 * ncp.value)
            w = np.zeros(npt.value)
            self._L.umcp_class_table(self._h, c, _dp(tab), _dp(w))
            out.append(dict(valence=val.value, ncp=ncp.value, npt=npt.value, nf=nf.value,
                            table=tab.reshape(npt.value, 6, ncp.value), weights=w))
        return out

    def set_launch(self, point_split: int, flags: int = 0):
        """point_split waves share 64 faces' Gauss points; flags: LAUNCH_*."""
        self._check(self._L.umcp_set_launch(self._h, point_split, int(flags)), "set_launch")

    def get_launch(self):
        sp, fl = C.c_int(), C.c_int()
        self._check(self._L.umcp_get_launch(self._h, C.byref(sp), C.byref(fl)), "get_launch")
        return sp.value, fl.value

    def profile_begin(self, every: int = 1):
        """Times the face launch of one evaluation in every `every` until profile_end."""
        self._check(self._L.umcp_profile_sample(self._h, int(every)), "profile_sample")
        self._check(self._L.umcp_profile_begin(self._h), "profile_begin")

    def profile_end(self, entries: int = 1):
        """(summed milliseconds, launches) of the face kernel since
        profile_begin; with entries=3 two lists: face kernel, vertex gather,
        restraint pre-pass and finalize."""
        ms = np.zeros(entries)
        cnt = (C.c_int * entries)()
        self._check(self._L.umcp_profile_end(self._h, entries, _dp(ms), cnt), "profile_end")
        if entries == 1:
            return float(ms[0]), int(cnt[0])
        return [float(x) for x in ms], [int(x) for x in cnt]

    def face_points(self, f: int):
        ncp = C.c_int()
        cp = (C.c_int * 32)()
        self._check(self._L.umcp_face_points(self._h, f, C.byref(ncp), cp), "face_points")
        return list(cp)[:ncp.value]

    def face_lipids(self, f: int):
        lo, li = C.c_double(), C.c_double()
        self._check(self._L.umcp_face_lipids(self._h, f, C.byref(lo), C.byref(li)), "face_lipids")
        return lo.value, li.value


def bd_coefficients(D_m2_per_s: float, dt_s: float, T_K: float = 298.0):
    """(mobility_dt, noise_amp) of r <- r - (D dt/kT) dE/dr + sqrt(2 D dt) xi."""
    D = units.diffusion_A2_per_s(D_m2_per_s)
    return D * dt_s / units.kT(T_K), float(np.sqrt(2.0 * D * dt_s))


def write_restart(path: str, r: np.ndarray, pp: np.ndarray | None = None, seed: int = 0):
    """Restart file of positions r (3nv+3) and optional momenta pp (3nv)."""
    L = load_library()
    r = np.ascontiguousarray(r, dtype=np.float64)
    nv = (r.size - 3) // 3
    p = None if pp is None else np.ascontiguousarray(pp, dtype=np.float64)
    rc = L.umcp_write_restart(path.encode(), nv, _dp(r), _dp(p) if p is not None else None, int(seed))
    if rc != 0:
        raise UmcpError("write_restart failed (%d)" % rc)


def read_restart(path: str, nv: int):
    """Returns dict(r, pp or None, alpha, seed) of a restart file of nv vertices."""
    L = load_library()
    r = np.zeros(3 * nv + 3)
    pp = np.zeros(3 * nv)
    al = np.zeros(3)
    has, seed = C.c_int(), C.c_int()
    rc = L.umcp_read_restart(path.encode(), nv, _dp(r), _dp(pp), _dp(al), C.byref(has), C.byref(seed))
    if rc != 0:
        raise UmcpError("read_restart failed (%d): %s" % (rc, path))
    return {"r": r, "pp": pp if has.value else None, "alpha": al, "seed": seed.value}


def noise(seed: int, step: int, n: int, d_out, stream=None):
    """The device noise of (seed, step) for vertices [0, n) into d_out (3n)."""
    L = load_library()
    rc = L.umcp_noise(C.c_ulonglong(seed), int(step), int(n), C.c_void_p(d_out.data_ptr()),
                      C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise UmcpError("noise failed (%d)" % rc)


class Ensemble:
    """Many resident surfaces stepped together in batched launches."""

    def __init__(self):
        self._L = load_library()
        h = C.c_void_p()
        rc = self._L.umcp_ensemble_create(C.byref(h))
        if rc != 0:
            raise UmcpError("ensemble_create failed (%d)" % rc)
        self._h = h
        self._members = {}

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise UmcpError("%s failed (%d): %s" % (what, rc, self._L.umcp_ensemble_last_error(self._h).decode()))

    def close(self):
        if self._h:
            self._L.umcp_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, surface: Surface, d_r, mobility_dt: float, noise_amp: float, seed: int, d_energies=None):
        bd = BDParams(mobility_dt, noise_amp)
        cap = int(d_energies.numel()) if d_energies is not None else 0
        rc = self._L.umcp_ensemble_add(self._h, surface._h, C.c_void_p(d_r.data_ptr()), C.byref(bd),
                                       C.c_ulonglong(seed),
                                       C.c_void_p(d_energies.data_ptr()) if d_energies is not None else None,
                                       cap)
        self._check(rc, "ensemble_add")
        self._members[id(surface)] = (surface, d_r, d_energies)   # keep the arrays alive

    def remove(self, surface: Surface):
        self._check(self._L.umcp_ensemble_remove(self._h, surface._h), "ensemble_remove")
        self._members.pop(id(surface), None)

    def __len__(self):
        return self._L.umcp_ensemble_size(self._h)

    def refresh(self):
        self._check(self._L.umcp_ensemble_refresh(self._h), "ensemble_refresh")

    def set_launch(self, point_split: int = 4, flags: int = LAUNCH_OCC2, face_grid: int = 0, order: int = 0):
        self._check(self._L.umcp_ensemble_set_launch(self._h, point_split, flags, face_grid, order),
                    "ensemble_set_launch")

    def get_launch(self):
        v = [C.c_int() for _ in range(4)]
        self._check(self._L.umcp_ensemble_get_launch(self._h, *[C.byref(x) for x in v]), "ensemble_get_launch")
        return tuple(x.value for x in v)

    def bd_steps(self, nsteps: int, graph_steps: int = 0, stream=None):
        self._check(self._L.umcp_ensemble_bd_steps(self._h, int(nsteps), int(graph_steps),
                                                   C.c_void_p(stream) if stream else None), "ensemble_bd_steps")

    def work(self):
        v = [C.c_int() for _ in range(3)]
        self._check(self._L.umcp_ensemble_work(self._h, *[C.byref(x) for x in v]), "ensemble_work")
        return tuple(x.value for x in v)

    def nan_check(self):
        """None if no member had a non-finite energy since the last check, else
        the first such member's index (every member's report is cleared)."""
        m = C.c_int(-1)
        rc = self._L.umcp_ensemble_nan_check(self._h, C.byref(m))
        if rc == -2:
            return m.value
        self._check(rc, "ensemble_nan_check")
        return None


class Server:
    """The serving loop: Brownian-dynamics requests admitted as they arrive
    and stepped together on the device (umcp_server_* in include/umcp.h)."""

    def __init__(self, budget_bytes: float = 200e9, max_active: int = 64, chunk_steps: int = 100,
                 precision: int = PREC_MIXED, point_split: int = 4, flags: int = LAUNCH_OCC2,
                 face_grid: int = 0, order: int = 0, graph_steps: int = 0, prep_threads: int = 4):
        self._L = load_library()
        cfg = ServerConfig(budget_bytes, max_active, chunk_steps, precision, point_split, flags, face_grid,
                           order, graph_steps, prep_threads)
        h = C.c_void_p()
        rc = self._L.umcp_server_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise UmcpError("server_create failed (%d)" % rc)
        self._h = h
        self._pending = {}

    def submit(self, mesh_path: str, nv: int, params: dict, c0: float, steps: int, seed: int,
               mobility_dt: float, noise_amp: float, r0: np.ndarray, r: np.ndarray) -> int:
        p = dict(kc=14.0, kg=0.0, KA=0.0, k_reg=0.0, mode=MODE_LIPID, precision=PREC_FP64, lipid_c0=0.0,
                 lipid_apl=68.3, grad_form=GRAD_REFERENCE)
        p.update(params)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        if r0.size != 3 * nv + 3 or r.size != 3 * nv + 3:
            raise UmcpError("positions must hold 3nv+3 doubles")
        rq = Request(mesh_path.encode(), nv,
                     Params(p["kc"], p["kg"], p["KA"], p["k_reg"], p["mode"], p["precision"], p["lipid_c0"],
                            p["lipid_apl"], p["grad_form"]), c0, int(steps), C.c_ulonglong(seed), BDParams(mobility_dt, noise_amp),
                     _dp(r0), _dp(r))
        rid = C.c_int()
        rc = self._L.umcp_server_submit(self._h, C.byref(rq), C.byref(rid))
        if rc != 0:
            raise UmcpError("server_submit failed (%d)" % rc)
        self._pending[rid.value] = (nv, int(steps))
        return rid.value

    def wait(self, rid: int):
        """(positions, energies per step, times [submitted, admitted, first step, done])."""
        nv, steps = self._pending.pop(rid)
        r = np.zeros(3 * nv + 3)
        e = np.zeros(max(steps, 1))
        t = np.zeros(4)
        rc = self._L.umcp_server_wait(self._h, rid, _dp(r), _dp(e), _dp(t))
        if rc != 0:
            raise UmcpError("request %d failed (%d): %s" % (rid, rc, self._L.umcp_server_last_error(self._h).decode()))
        return r, e[:steps], t

    def drain(self):
        self._L.umcp_server_drain(self._h)

    def stats(self) -> dict:
        v = np.zeros(12)
        self._L.umcp_server_stats(self._h, _dp(v), len(v))
        keys = ("done", "chunks", "request_steps", "face_steps", "busy_s", "peak_bytes", "peak_active", "clock_s",
                "build_s", "retire_s", "wait_s", "setup_s")
        return dict(zip(keys, v.tolist()))

    def close(self):
        if self._h:
            self._L.umcp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
