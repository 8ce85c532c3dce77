"""Internal units (kpc, Myr, M_sun) and the gravitational constant in them.

G is derived from the standard astronomical value 4.30091e-3 pc (km/s)^2 / M_sun with the
IAU parsec and the Julian year, giving G ~ 4.498e-12 kpc^3 M_sun^-1 Myr^-2 — the same unit
system as the reference (units.py). The PM configurations used for benchmarking run in
G = 1 units instead and pass ``grav_constant`` explicitly.
"""

from __future__ import annotations

METRES_PER_PARSEC = 3.0856775814913673e16
KM_PER_KPC = METRES_PER_PARSEC * 1.0e3 * 1.0e-3
SECONDS_PER_MYR = 365.25 * 86400.0 * 1.0e6
KMS_PER_KPC_PER_MYR = KM_PER_KPC / SECONDS_PER_MYR  # ~977.79 km/s per kpc/Myr
G_PC_KMS2 = 4.30091e-3
G = G_PC_KMS2 * 1.0e-3 / KMS_PER_KPC_PER_MYR ** 2


def kms_to_kpc_per_myr(v_kms: float) -> float:
    return v_kms / KMS_PER_KPC_PER_MYR


def kpc_per_myr_to_kms(v: float) -> float:
    return v * KMS_PER_KPC_PER_MYR


__all__ = ["G", "KMS_PER_KPC_PER_MYR", "kms_to_kpc_per_myr", "kpc_per_myr_to_kms"]
