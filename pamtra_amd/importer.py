"""Profile builders (python/pyPamtra/importer.py).  Only the clear-sky U.S.
standard atmosphere helper (importer.py:3039-3095) is provided; the NWP
model-file readers are outside this library."""

from __future__ import annotations

import numpy as np

from .core import pyPamtra
from .meteoSI import usStandard


def _createUsStandardProfile(**kwargs):
    """Level pressure and temperature of the 1976 standard atmosphere at
    hgt_lev (mandatory); humidity zero unless given."""
    assert "hgt_lev" in kwargs
    hgt = np.asarray(kwargs["hgt_lev"], dtype=np.float64)
    pamData = dict()
    _, p, t = usStandard(hgt)
    pamData["press_lev"] = p
    pamData["temp_lev"] = t
    for k, v in kwargs.items():
        pamData[k] = np.array(v)
    if "relhum_lev" not in kwargs and "q" not in kwargs:
        pamData["relhum_lev"] = np.zeros_like(hgt)
    return pamData


def createUsStandardProfile(pam=None, **kwargs):
    """Fills `pam` with a clear-sky U.S. standard atmosphere at hgt_lev."""
    if pam is None:
        pam = pyPamtra()
    pam.createProfile(**_createUsStandardProfile(**kwargs))
    return pam
