"""Per-trial scalar set-up of the strength-reduction loop (host side).

These run once per SRF on a handful of material blocks, so they stay on the
host, evaluated in the reference's operation order with the same libm calls
(math.sin/cos/tan/atan are the C library's).
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

PI = 3.141592653589793               # src/global.f90:19
DEG2RAD = PI / 180.0                 # src/global.f90:42
RAD2DEG = 180.0 / PI
ZEROTOL = 1.0e-12                    # src/global.f90:22


def strength_reduction(srf: float, phinu: bool, coh, nu, phi, psi
                       ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, int]:
    """src/plastic_library.f90:42-88.  Returns (cohf, nuf, phif, psif, istat);
    istat = 1 when the phi-nu correction changed Poisson's ratio."""
    cohf = np.array(coh, dtype=np.float64)
    nuf = np.array(nu, dtype=np.float64)
    phif = np.array(phi, dtype=np.float64)
    psif = np.array(psi, dtype=np.float64)
    istat = 0
    if srf != 1.0:
        for m in range(len(cohf)):
            tnphi = math.tan(phi[m] * DEG2RAD)
            phif[m] = math.atan(tnphi / srf) * RAD2DEG
            tnpsi = math.tan(psi[m] * DEG2RAD)
            psif[m] = math.atan(tnpsi / srf) * RAD2DEG
            cohf[m] = coh[m] / srf
    if not phinu:
        return cohf, nuf, phif, psif, istat
    for m in range(len(cohf)):
        omtnu = 1.0 - 2.0 * nu[m]
        snphif = math.sin(phif[m] * DEG2RAD)
        if snphif < omtnu:
            snphi = math.sin(phi[m] * DEG2RAD)
            beta = snphi / omtnu
            if beta < 1.0:
                beta = 1.0 + ZEROTOL
            nuf[m] = 0.5 * (1.0 - snphif / beta)
            istat = 1
    return cohf, nuf, phif, psif, istat


def dt_viscoplas(nuf: Sequence[float], phif: Sequence[float],
                 ymf: Sequence[float], ismat=None) -> float:
    """Stable pseudo-time step, src/plastic_library.f90:17-39."""
    dt_min = 1.0e32
    for m in range(len(nuf)):
        if ismat is not None and not ismat[m]:
            continue
        snphi = math.sin(phif[m] * DEG2RAD)
        dt = (4.0 * (1.0 + nuf[m]) * (1.0 - 2.0 * nuf[m])
              / (ymf[m] * (1.0 - 2.0 * nuf[m] + snphi * snphi)))
        if dt < dt_min:
            dt_min = dt
    return dt_min
