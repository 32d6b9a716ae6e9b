"""The reference's golden rain/snow scenario (reference tests/_scenario.py:14-44)
built through this package's pyPamtra API."""

import numpy as np

import pamtra_amd
from pamtra_amd import importer

FREQUENCIES = [10.65, 35.5]


def build_pamtra(geopotential=True):
    pam = pamtra_amd.pyPamtra()
    pam.df.addHydrometeor((
        "rwc_q", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0,
        3, 30, "exp", -99.0, -99.0, 8.0e6, -99.0, 1.0e-5, 6.0e-3,
        "mie-sphere", "khvorostyanov01_drops", -99.0,
    ))
    pam.df.addHydrometeor((
        "swc_q", -99.0, -1, -99.0, 0.038, 2.0, 0.3971, 1.88,
        3, 30, "exp_field_t", -99.0, -99.0, -99.0, -99.0, 0.51e-10, 2.0e-2,
        "mie-sphere", "heymsfield10_particles", -99.0,
    ))
    hl = np.array([0.0, 1000.0, 2000.0, 3000.0])
    if geopotential:
        pam = importer.createUsStandardProfile(pam, hgt_lev=hl)
    else:
        _, p, t = pamtra_amd.meteoSI.usStandard(hl, geopotential=False)
        pam.createProfile(hgt_lev=hl, press_lev=p, temp_lev=t, relhum_lev=np.zeros(4))
    pam.set["verbose"] = 0
    pam.set["pyVerbose"] = 0
    pam.nmlSet["randomseed"] = 10
    pam.p["hydro_q"][0, 0, 1, 0] = 1e-3
    pam.p["hydro_q"][0, 0, 2, 1] = 5e-4
    return pam


def passive_nml(pam):
    nml = dict(pam.nmlSet)
    nml["active"] = False
    return nml
