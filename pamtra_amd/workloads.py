"""Seeded synthetic workloads of the benchmark configurations (passive
Mie + RT4).  Each builder returns a ready ``pyPamtra`` object (descriptor
table, profile, nmlSet) and the channel list; the arrays can be moved to the
device for ``pam_run_device``.

  config1: 64 columns x 30 layers (0-12 km), one channel (89 GHz), cloud
           (modified gamma) + rain (Marshall-Palmer), 50 bins, ocean.
  config2: 16384 columns, 20-120 layers, tops 10-20 km, 7 channels,
           cloud + rain with 80 bins, ocean (TESSEM2) or Lambertian land.
  config3: the config2 shapes with four categories of 100 bins: cloud
           liquid (modified gamma), cloud ice (917 kg/m3 spheres), snow (soft
           spheres, 50-300 kg/m3 per column) and graupel (400 kg/m3).
  serving_requests: the serving mix, requests of 1-4096 columns with
           20-150 layers, 1-12 channels and 0-4 of the config3 categories.
  config5: the large domain (1024 x 1024 columns in full) of 150 layers, 12
           channels 10.65-325 GHz, the config3 categories, 32 quadrature
           angles per hemisphere (64x64 operators); built for a column
           sample of any size.
  tmatrix_config: the config1 grid (30 layers, 0-12 km, ocean) with cloud
           droplets as T-matrix spheres and rain as oblate T-matrix
           spheroids (as_ratio 0.7), 30 bins each, 3 channels 22.235-89 GHz.
Random draws use numpy.random.default_rng(seed), fields in a fixed order.
"""

from __future__ import annotations

import numpy as np

from .core import pyPamtra

R_D = 287.0596736665907
CONFIG2_FREQS = [22.235, 23.84, 31.4, 50.3, 89.0, 150.0, 183.31]


def _temperature(h):
    """288 K at the ground, -6.5 K/km up to 11 km, isothermal above."""
    return 288.0 - 6.5e-3 * np.minimum(h, 11000.0)


def _hydrostatic_pressure(h_lev, t_lev, p0=101325.0, g=9.80665):
    """Pressure at the levels from the hypsometric equation, layer-mean T."""
    p = np.empty_like(h_lev)
    p[..., 0] = p0
    for k in range(1, h_lev.shape[-1]):
        tm = 0.5 * (t_lev[..., k] + t_lev[..., k - 1])
        dz = h_lev[..., k] - h_lev[..., k - 1]
        p[..., k] = p[..., k - 1] * np.exp(-g * dz / (R_D * tm))
    return p


def _descriptor(pam, nbin, d1=1e-6, d2=6e-3):
    # cloud: modified gamma, mu = 2, gamma = 1, N_tot = 1e8 m^-3 via p_1
    pam.df.addHydrometeor(("cwc_q", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "mgamma", 1.0e8, -99.0, 2.0, 1.0, d1, d2, "mie-sphere",
                           "khvorostyanov01_drops", -99.0))
    # rain: exponential, N0 = 8e6 m^-4 via p_3
    pam.df.addHydrometeor(("rwc_q", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "exp", -99.0, -99.0, 8.0e6, -99.0, d1, d2, "mie-sphere",
                           "khvorostyanov01_drops", -99.0))


def _common_nml(pam):
    pam.nmlSet["active"] = False
    pam.nmlSet["passive"] = True
    pam.nmlSet["hydro_adaptive_grid"] = False
    pam.nmlSet["conserve_mass_rescale_dsd"] = True
    pam.nmlSet["hydro_includehydroinrhoair"] = True
    pam.nmlSet["gas_mod"] = "R98"
    pam.nmlSet["liq_mod"] = "TKC"
    pam.nmlSet["outpol"] = "VH"


def _columns(rng, ncol, nlyrs, tops, max_nlyrs):
    """Level heights, temperature, pressure, relative humidity (padded)."""
    hgt_lev = np.full((ncol, max_nlyrs + 1), np.nan)
    for c in range(ncol):
        hgt_lev[c, :nlyrs[c] + 1] = np.linspace(0.0, tops[c], nlyrs[c] + 1)
    # pad with a continued grid so that every array stays finite
    for c in range(ncol):
        L = nlyrs[c]
        if L < max_nlyrs:
            dz = tops[c] / L
            hgt_lev[c, L + 1:] = tops[c] + dz * np.arange(1, max_nlyrs - L + 1)
    t_lev = _temperature(hgt_lev)
    p_lev = _hydrostatic_pressure(hgt_lev, t_lev)
    rh_lev = rng.uniform(20.0, 100.0, size=(ncol, max_nlyrs + 1))
    return hgt_lev, t_lev, p_lev, rh_lev


def _hydro_q(rng, hgt, t, p, ncol, max_nlyrs, nlyrs):
    """Cloud LWC log-uniform 1e-3..0.5 g/m3 in 1-3 km, rain from
    Marshall-Palmer R ~ U(0, 10) mm/h below 2 km; returned in kg/kg."""
    rho = p / (R_D * t)
    q = np.zeros((ncol, max_nlyrs, 2))
    lwc = 10.0 ** rng.uniform(-3.0, np.log10(0.5), size=(ncol, max_nlyrs)) * 1e-3
    rr = rng.uniform(0.0, 10.0, size=ncol)
    lam = 4100.0 * np.maximum(rr, 1e-3) ** -0.21
    w_rain = np.pi * 1000.0 * 8.0e6 * lam ** -4.0
    valid = np.arange(max_nlyrs)[None, :] < nlyrs[:, None]
    cloud = (hgt >= 1000.0) & (hgt <= 3000.0) & valid
    rain = (hgt < 2000.0) & valid & (rr[:, None] > 0.05)
    q[..., 0] = np.where(cloud, lwc / rho, 0.0)
    q[..., 1] = np.where(rain, w_rain[:, None] / rho, 0.0)
    return q


def config1(ncol=64, seed=1001, nbin=50):
    """Plain FP64 passive RT4+Mie, 89 GHz, ocean (TESSEM2)."""
    rng = np.random.default_rng(seed)
    pam = pyPamtra()
    _descriptor(pam, nbin)
    L = 30
    nlyrs = np.full(ncol, L)
    tops = np.full(ncol, 12000.0)
    hgt_lev, t_lev, p_lev, rh_lev = _columns(rng, ncol, nlyrs, tops, L)
    hgt = 0.5 * (hgt_lev[:, 1:] + hgt_lev[:, :-1])
    t = 0.5 * (t_lev[:, 1:] + t_lev[:, :-1])
    p = np.exp(0.5 * (np.log(p_lev[:, 1:]) + np.log(p_lev[:, :-1])))
    q = _hydro_q(rng, hgt, t, p, ncol, L, nlyrs)
    sst = rng.uniform(280.0, 300.0, size=ncol)
    pam.createProfile(hgt_lev=hgt_lev, temp_lev=t_lev, press_lev=p_lev, relhum_lev=rh_lev,
                      hydro_q=q, groundtemp=sst, sfc_type=np.zeros(ncol, dtype=int),
                      sfc_model=np.zeros(ncol, dtype=int), sfc_refl=np.full(ncol, "S"),
                      sfc_salinity=np.full(ncol, 33.0), wind10u=np.zeros(ncol),
                      wind10v=np.zeros(ncol), timestamp=0)
    _common_nml(pam)
    return pam, [89.0]


TMATRIX_FREQS = [22.235, 35.5, 89.0]


def tmatrix_config(ncol=64, seed=1006, nbin=30, freqs=None):
    """Passive RT4 with the T-matrix producer: the config1 grid and
    hydrometeor fields, cloud droplets as spheres and rain as oblate
    spheroids (axis ratio 0.7), canting 0."""
    pam, _ = config1(ncol=ncol, seed=seed, nbin=nbin)
    pam.df.data["scat_name"][:] = "tmatrix"
    pam.df.data["as_ratio"][0] = 1.0
    pam.df.data["as_ratio"][1] = 0.7
    pam.df.data["canting"][:] = 0.0
    return pam, list(freqs or TMATRIX_FREQS)


def config2(ncol=16384, seed=1002, nbin=80, freqs=None):
    """Batched FP64: 20-120 layers, tops 10-20 km, 7 channels, ocean or land."""
    rng = np.random.default_rng(seed)
    pam = pyPamtra()
    _descriptor(pam, nbin)
    ML = 120
    nlyrs = rng.integers(20, 121, size=ncol)
    tops = rng.uniform(10000.0, 20000.0, size=ncol)
    hgt_lev, t_lev, p_lev, rh_lev = _columns(rng, ncol, nlyrs, tops, ML)
    hgt = 0.5 * (hgt_lev[:, 1:] + hgt_lev[:, :-1])
    t = 0.5 * (t_lev[:, 1:] + t_lev[:, :-1])
    p = np.exp(0.5 * (np.log(p_lev[:, 1:]) + np.log(p_lev[:, :-1])))
    q = _hydro_q(rng, hgt, t, p, ncol, ML, nlyrs)
    land = rng.uniform(size=ncol) < 0.5
    eps = np.round(rng.uniform(0.85, 0.95, size=ncol), 2)
    sst = rng.uniform(280.0, 300.0, size=ncol)
    hgt_pad = hgt.copy()
    for c in range(ncol):
        hgt_pad[c, nlyrs[c]:] = -9999.0
    pam.createProfile(hgt_lev=hgt_lev, hgt=hgt_pad, temp_lev=t_lev, press_lev=p_lev,
                      relhum_lev=rh_lev, hydro_q=q, groundtemp=sst,
                      sfc_type=np.where(land, -9999, 0), sfc_model=np.where(land, -9999, 0),
                      sfc_refl=np.where(land, "L", "S"), sfc_emissivity=eps,
                      sfc_salinity=np.full(ncol, 33.0), wind10u=np.zeros(ncol),
                      wind10v=np.zeros(ncol), timestamp=0)
    # layer heights beyond nlyrs are missing; use the level midpoints there
    pam.p["hgt"] = np.where(pam.p["hgt"] == -9999.0,
                            0.5 * (hgt_lev[:, 1:] + hgt_lev[:, :-1])[:, None, :],
                            pam.p["hgt"]) if pam.p["hgt"].ndim == 3 else pam.p["hgt"]
    _common_nml(pam)
    return pam, list(CONFIG2_FREQS if freqs is None else freqs)


CONFIG3_NBIN = 100
SERVING_FREQS = [10.65, 18.7, 22.235, 23.8, 31.4, 36.5, 50.3, 89.0, 150.0, 166.0, 183.31, 325.0]


def _descriptor4(pam, nbin):
    """Cloud liquid, cloud ice, snow, graupel (mass from the density, so the
    exponential spectra take N0 via p_3).  Large particles reach 13 mm, a
    size parameter of ~25 at 183.31 GHz."""
    pam.df.addHydrometeor(("cwc_q", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "mgamma", 1.0e8, -99.0, 2.0, 1.0, 1e-6, 6e-3, "mie-sphere",
                           "khvorostyanov01_drops", -99.0))
    pam.df.addHydrometeor(("iwc_q", -99.0, -1, 917.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "exp", -99.0, -99.0, 1.0e7, -99.0, 1e-6, 1.3e-2, "mie-sphere",
                           "heymsfield10_particles", -99.0))
    pam.df.addHydrometeor(("swc_q", -99.0, -1, 100.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "exp", -99.0, -99.0, 3.0e6, -99.0, 1e-5, 1.3e-2, "mie-sphere",
                           "heymsfield10_particles", -99.0))
    pam.df.addHydrometeor(("gwc_q", -99.0, -1, 400.0, -99.0, -99.0, -99.0, -99.0, 3, nbin,
                           "exp", -99.0, -99.0, 4.0e6, -99.0, 1e-5, 1.3e-2, "mie-sphere",
                           "heymsfield10_particles", -99.0))


def _hydro_q4(rng, hgt, t, p, ncol, max_nlyrs, nlyrs, present=None):
    """Contents (kg/kg) of the four categories: cloud LWC log-U(1e-3, 0.5)
    g/m3 in 1-3 km; ice IWC log-U(1e-4, 0.2) g/m3 in 6-11 km; snow log-U(1e-3,
    0.5) g/m3 in 3-8 km; graupel log-U(1e-3, 0.5) g/m3 in 2-6 km.  `present`
    (ncol, 4) switches categories off per column."""
    rho = p / (R_D * t)
    valid = np.arange(max_nlyrs)[None, :] < nlyrs[:, None]
    spec = [(-3.0, np.log10(0.5), 1000.0, 3000.0), (-4.0, np.log10(0.2), 6000.0, 11000.0),
            (-3.0, np.log10(0.5), 3000.0, 8000.0), (-3.0, np.log10(0.5), 2000.0, 6000.0)]
    q = np.zeros((ncol, max_nlyrs, 4))
    for h, (lo, hi, z0, z1) in enumerate(spec):
        wc = 10.0 ** rng.uniform(lo, hi, size=(ncol, max_nlyrs)) * 1e-3
        inside = (hgt >= z0) & (hgt <= z1) & valid
        if present is not None:
            inside &= present[:, h][:, None]
        q[..., h] = np.where(inside, wc / rho, 0.0)
    return q


def _profile_columns(rng, ncol, nlyrs, tops, ML, nbin, present=None):
    pam = pyPamtra()
    _descriptor4(pam, nbin)
    hgt_lev, t_lev, p_lev, rh_lev = _columns(rng, ncol, nlyrs, tops, ML)
    hgt = 0.5 * (hgt_lev[:, 1:] + hgt_lev[:, :-1])
    t = 0.5 * (t_lev[:, 1:] + t_lev[:, :-1])
    p = np.exp(0.5 * (np.log(p_lev[:, 1:]) + np.log(p_lev[:, :-1])))
    q = _hydro_q4(rng, hgt, t, p, ncol, ML, nlyrs, present)
    land = rng.uniform(size=ncol) < 0.5
    eps = np.round(rng.uniform(0.85, 0.95, size=ncol), 2)
    sst = rng.uniform(280.0, 300.0, size=ncol)
    rho_snow = rng.uniform(50.0, 300.0, size=ncol)
    pam.createProfile(hgt_lev=hgt_lev, temp_lev=t_lev, press_lev=p_lev, relhum_lev=rh_lev,
                      hydro_q=q, groundtemp=sst,
                      sfc_type=np.where(land, -9999, 0), sfc_model=np.where(land, -9999, 0),
                      sfc_refl=np.where(land, "L", "S"), sfc_emissivity=eps,
                      sfc_salinity=np.full(ncol, 33.0), wind10u=np.zeros(ncol),
                      wind10v=np.zeros(ncol), timestamp=0)
    # per-column snow density (soft spheres) through the 4D descriptor
    rho_ms = np.empty((ncol, 1, ML, 4))
    rho_ms[..., 0] = -99.0
    rho_ms[..., 1] = 917.0
    rho_ms[..., 2] = rho_snow[:, None, None]
    rho_ms[..., 3] = 400.0
    pam.df.add4D("rho_ms", rho_ms)
    _common_nml(pam)
    return pam


def config3(ncol=16384, seed=1003, nbin=CONFIG3_NBIN, freqs=None):
    """Mixed-precision shapes: config2 grid and channels, four categories."""
    rng = np.random.default_rng(seed)
    ML = 120
    nlyrs = rng.integers(20, 121, size=ncol)
    tops = rng.uniform(10000.0, 20000.0, size=ncol)
    pam = _profile_columns(rng, ncol, nlyrs, tops, ML, nbin)
    return pam, list(CONFIG2_FREQS if freqs is None else freqs)


def serving_requests(n, seed=1004, max_cols=4096, nbin=CONFIG3_NBIN):
    """n requests of the serving mix: columns log-uniform in [1, max_cols],
    20-150 layers per column, 1-12 channels drawn from SERVING_FREQS, 0-4
    categories.  Returns a list of (pyPamtra, freqs)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ncol = int(np.clip(np.round(np.exp(rng.uniform(0.0, np.log(max_cols)))), 1, max_cols))
        nlyrs = rng.integers(20, 151, size=ncol)
        ML = int(nlyrs.max())
        tops = rng.uniform(10000.0, 20000.0, size=ncol)
        nch = int(rng.integers(1, 13))
        freqs = sorted(rng.choice(SERVING_FREQS, size=nch, replace=False).tolist())
        ncat = int(rng.integers(0, 5))
        cats = rng.choice(4, size=ncat, replace=False)
        present = np.zeros((ncol, 4), dtype=bool)
        present[:, cats] = True
        pam = _profile_columns(rng, ncol, nlyrs, tops, ML, nbin, present)
        out.append((pam, freqs))
    return out


def config5(ncol=1024 * 1024, seed=1005, nbin=CONFIG3_NBIN, freqs=None):
    """Large domain: 150 layers, tops 10-20 km, 12 channels, four categories,
    32 angles per hemisphere (pyPamtra's _nangles = 32)."""
    rng = np.random.default_rng(seed)
    ML = 150
    nlyrs = np.full(ncol, ML)
    tops = rng.uniform(10000.0, 20000.0, size=ncol)
    pam = _profile_columns(rng, ncol, nlyrs, tops, ML, nbin)
    pam._nangles = 32
    return pam, list(SERVING_FREQS if freqs is None else freqs)
