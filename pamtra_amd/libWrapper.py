"""The boundary between the pyPamtra dictionaries and the MI355X library:
the replacement of python/pyPamtra/libWrapper.py:24-235.

``PamtraFortranWrapper`` keeps the reference's signature and return tuple,
but instead of filling f2py module globals it marshals the dictionaries into
the C-ABI of include/pamtra_amd.h and calls ``pam_run`` of
``libpamtra_amd.so`` through ctypes.  There is no CPU fallback: without the
HIP library or without a GPU the call raises.
"""

from __future__ import annotations

import copy
import ctypes as C
import warnings

import numpy as np

from . import _capi
from ._lib import get_context, library

# radar outputs of the "simple" mode (results key -> pam_results array)
RADAR_KEYS = {"Ze": "ze", "Att_hydro": "att_hydro", "Att_atmo": "att_atmo",
              "radar_hgt": "radar_hgt"}
RESULT_KEYS = ["tb", "Ze", "emissivity", "Att_hydro", "Att_atmo", "radar_hgt", "radar_moments",
               "radar_edges", "radar_slopes", "radar_quality", "radar_snr", "radar_spectra",
               "radar_vel", "psd_d", "psd_deltad", "psd_n", "psd_mass", "psd_area", "psd_bscat",
               "kextatmo", "scatter_matrix", "extinct_matrix", "emis_vector", "angles_deg"]


class _ModuleHandle(object):
    """Stand-in for the f2py module object the reference returns
    (libWrapper.py:231-233): exposes the library version."""

    def __init__(self, lib):
        self.version = lib.pam_version().decode()


def PamtraFortranWrapper(sets, nmlSets, descriptorFile, descriptorFile4D, descriptorFileFS,
                         profile, returnModule=True):
    """Runs the passive forward operator on the GPU.  Returns
    ``(results, pamError[, module])`` like libWrapper.py:24-235."""
    for key, v in profile.items():
        if isinstance(v, np.ndarray) and v.ndim >= 2 and "lat" in profile:
            assert v.shape[0] == profile["lat"].shape[0]
            assert v.shape[1] == profile["lat"].shape[1]
    nml = dict(nmlSets)
    if not nml.get("passive", True):
        raise NotImplementedError("passive=False leaves nothing for this library to compute")
    m = _capi.marshal(nml, sets, descriptorFile, descriptorFile4D, profile, descriptorFileFS)
    lib = library(m.nummu)
    ctx = get_context(0, m.nummu)
    freqs, fp = _capi.freqs_array(sets["freqs"])
    hr = _capi.alloc_results(m, len(freqs), diagnostics=bool(nml.get("save_ssp", False)))
    err = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                      C.byref(m.profile), fp, len(freqs), C.byref(hr.struct))
    if err == _capi.PAM_FATAL and sets.get("verbose", 0) >= 0:
        msg = _capi.results_message(hr)
        if msg:
            warnings.warn("pam_run: " + msg, RuntimeWarning)
    results = dict()
    for key in RESULT_KEYS:
        if key in ("tb", "emissivity", "angles_deg"):
            results[key] = copy.deepcopy(hr.arrays[key])
        elif key == "kextatmo" and "kextatmo" in hr.arrays:
            results[key] = copy.deepcopy(hr.arrays[key])
        elif key in RADAR_KEYS and RADAR_KEYS[key] in hr.arrays:
            results[key] = copy.deepcopy(hr.arrays[RADAR_KEYS[key]])
        elif key == "radar_quality":
            results[key] = -9999
        else:
            results[key] = -9999.
    results["pamtraVersion"] = lib.pam_version().decode()
    results["pamtraHash"] = "pamtra_amd"
    results["status"] = hr.arrays["status"]
    if returnModule:
        return results, int(err), _ModuleHandle(lib)
    return results, int(err)
