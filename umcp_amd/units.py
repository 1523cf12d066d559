"""Unit conventions of the reference, frozen in one place.

Energies in kcal/mol, lengths in Angstrom, time in seconds at the boundary.
kT = 0.5922 / 298 * T kcal/mol (src/hd.C:143).  The internal AKMA time unit
(src/units.h:5) is only needed by the momentum-space integrator.
"""
from __future__ import annotations

KT_PER_K = 0.5922 / 298.0          # kcal/mol per kelvin (src/hd.C:143)
AKMA_TIME = 1.0 / 0.04888e-12      # 1/s (src/units.h:5)
M2_TO_A2 = 1e20


def kT(T_K: float = 298.0) -> float:
    return KT_PER_K * T_K


def diffusion_A2_per_s(D_m2_per_s: float) -> float:
    return D_m2_per_s * M2_TO_A2


def kc_from_kT(kc_in_kT: float, T_K: float = 298.0) -> float:
    return kc_in_kT * kT(T_K)


def KA_from_kT(KA_in_kT_per_A2: float, T_K: float = 298.0) -> float:
    return KA_in_kT_per_A2 * kT(T_K)
