"""Thermodynamic helpers used to build test profiles.

The reference delegates ``usStandard`` to the unvendored ``meteo_si`` package
(python/pyPamtra/meteoSI.py:438).  This is a restatement of the U.S. Standard
Atmosphere 1976 (NOAA/NASA/USAF, 1976) below 86 km: piecewise-linear
temperature in geopotential height, hydrostatic pressure.
"""

from __future__ import annotations

import numpy as np

Grav = 9.80665          # standard gravity [m/s2]
Rstar = 8.31432         # universal gas constant of the 1976 standard [J/(mol K)]
M0 = 0.0289644          # molar mass of dry air [kg/mol]
Rair = Rstar / M0       # [J/(kg K)]
R_D = 287.0596736665907  # dry air [J/(kg K)] (constants.f90:29)
R_V = 461.5249933083879  # water vapour [J/(kg K)] (constants.f90:30)
R_EARTH = 6356766.0     # effective earth radius of the 1976 standard [m]
missingNumber = -9999.

_HB = np.array([0.0, 11000.0, 20000.0, 32000.0, 47000.0, 51000.0, 71000.0, 84852.0])
_LB = np.array([-6.5e-3, 0.0, 1.0e-3, 2.8e-3, 0.0, -2.8e-3, -2.0e-3, 0.0])
_T0, _P0 = 288.15, 101325.0


def _base_values():
    tb, pb = [_T0], [_P0]
    for i in range(len(_HB) - 1):
        dh = _HB[i + 1] - _HB[i]
        t1 = tb[-1] + _LB[i] * dh
        if _LB[i] == 0.0:
            p1 = pb[-1] * np.exp(-Grav * M0 * dh / (Rstar * tb[-1]))
        else:
            p1 = pb[-1] * (tb[-1] / t1) ** (Grav * M0 / (Rstar * _LB[i]))
        tb.append(t1)
        pb.append(p1)
    return np.array(tb), np.array(pb)


_TB, _PB = _base_values()


def usStandard(height, geopotential=True):
    """Density [kg/m3], pressure [Pa] and temperature [K] of the 1976 U.S.
    standard atmosphere at ``height`` [m].  With ``geopotential`` the
    geometric height is converted to geopotential height first."""
    h = np.asarray(height, dtype=np.float64)
    H = R_EARTH * h / (R_EARTH + h) if geopotential else h
    i = np.clip(np.searchsorted(_HB, H, side="right") - 1, 0, len(_HB) - 1)
    dh = H - _HB[i]
    T = _TB[i] + _LB[i] * dh
    iso = _LB[i] == 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        p_grad = _PB[i] * (_TB[i] / T) ** (Grav * M0 / (Rstar * np.where(iso, 1.0, _LB[i])))
        p_iso = _PB[i] * np.exp(-Grav * M0 * dh / (Rstar * _TB[i]))
    p = np.where(iso, p_iso, p_grad)
    rho = p / (Rair * T)
    return rho, p, T


def e_sat_gg_water(T):
    """Goff-Gratch saturation pressure over water [Pa]."""
    T = np.asarray(T, dtype=np.float64)
    s = (-7.90298 * (373.16 / T - 1) + 5.02808 * np.log10(373.16 / T)
         - 1.3816e-7 * (10 ** (11.344 * (1 - T / 373.16)) - 1)
         + 8.1328e-3 * (10 ** (-3.49149 * (373.16 / T - 1)) - 1))
    return 1013.246 * 10 ** s * 100.0


def rh2q(rh, T, p):
    """Specific humidity [kg/kg] from relative humidity [1], T [K], p [Pa]."""
    e = rh * e_sat_gg_water(T)
    return 0.622 * e / (p - 0.378 * e)


def moist_rho_rh(p, T, rh, *qm):
    """Density of moist air [kg/m3] from p [Pa], T [K], rh [1]; qm: contents
    [kg/kg] of other species in the air mass (meteoSI.py:51-73)."""
    if np.any(np.asarray(rh) > 5):
        raise TypeError("rh must not be in %")
    return moist_rho_q(p, T, rh2q(rh, T, p), *qm)


def moist_rho_q(p, T, q, *qm):
    """Density of moist air [kg/m3] from p [Pa], T [K], specific humidity q
    and the contents qm of other species: p / (R_d (1 + (R_v/R_d - 1) q - qm) T),
    the virtual-temperature form of conversions.f90:5-24."""
    qs = 0.0
    if len(qm) > 0:
        qs = np.ma.array(qm).filled(0)
        qs[qs < 0] = 0
        qs = np.sum(qs, axis=0)
    return p / (R_D * (1 + (R_V / R_D - 1) * q - qs) * T)


def detect_liq_cloud(z, t, rh):
    """Liquid cloud layers on a level grid (meteoSI.py:253-314): a layer is
    cloudy when the mean of its two levels' rh exceeds 0.95 and their mean
    temperature 253.15 K.  Returns (i_top, i_base, i_cloud): levels that only
    close a cloudy layer from above, levels that only open one, and all
    levels touching a cloudy layer; empty lists when tops and bases do not
    pair up."""
    n = len(z)
    flag = np.zeros(n, dtype=int)
    for i in range(n - 1):
        if (rh[i + 1] + rh[i]) / 2. > 0.95 and (t[i + 1] + t[i]) / 2. > 253.15:
            flag[i] |= 1
            flag[i + 1] |= 2
    i_cloud = np.where(flag != 0)[0]
    i_top, i_base = [], []
    if len(i_cloud):
        i_base = np.where(flag == 1)[0]
        i_top = np.where(flag == 2)[0]
        if len(i_top) != len(i_base):
            return [], [], []
    return i_top, i_base, i_cloud
